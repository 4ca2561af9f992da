"""Many FT-CCSD instances on one GPU: the serving-style workload.

A batch is a list of (T, mu, system) instances of uneven size n and grid
ng, each an ordinary ``ccsd(...).run()`` (the reference solves them one
after another; kelvin/ccsd.py has no batching).  ``solve_batch`` runs them
concurrently: each of at most four worker threads owns a HIP stream (the
device exposes four hardware queues to a process), and each instance runs
there with every fixed-point step replayed as a HIP graph
(``ccsd(graph=True)``), so the GPU overlaps the small instances' kernels and
the host issues one graph launch per iteration instead of a few hundred
kernel launches.  Instances are started shortest first (by the n^6 ng cost
of an iteration), which minimises the mean completion time.

``autotune`` chooses, per (n, ng, precision) class, the launch shapes of
the residual by timing it: first the number of grid points whose products
go out as one launch (the chunk), then the GEMM launch plan at that chunk
(tile shape policy and split-K threshold, ``ftcc_set_gemm_plan``).
``solve_batch`` applies the choices.

``python -m kelvin_amd.batch`` runs the 64-instance mix (n in {38, 54, 66},
i.e. 19, 27 or 33 k-points, ng in {8, 12, 16, 24}, T/T_F in [0.05, 1],
rs in [1, 5], N = 14, drawn from ``numpy.random.default_rng(4)``) both as
the plain loop and batched, and prints one JSON line with aggregate
instances/s, p50/p99 of each instance's own solve time and of its latency
(completion time from the start of the batch, every instance submitted at
once), and the largest difference in Omega between the two.
"""
import argparse
import concurrent.futures
import json
import os
import sys
import time

import numpy
import torch

from .ccsd import ccsd
from .engine import FTEngine
from . import quadrature
from .ueg_system import ueg_from_config

_NK = (19, 27, 33)
_NG = (8, 12, 16, 24)


def c4_instances(count=64, seed=4):
    """The serving mix: dicts (nk, ng, t_over_tf, rs, N)."""
    rng = numpy.random.default_rng(seed)
    out = []
    for _ in range(count):
        out.append(dict(nk=int(rng.choice(_NK)), ng=int(rng.choice(_NG)),
                        t_over_tf=float(rng.uniform(0.05, 1.0)),
                        rs=float(rng.uniform(1.0, 5.0)), N=14))
    return out


# GEMM launch plans the autotuner tries: (tile policy, split-K target)
PLANS = ((0, 1024), (1, 1024), (2, 1024), (0, 512), (0, 2048))


def autotune(nk, ng, precision="fp64", reps=3, rs=2.0, t_over_tf=0.5,
             plans=PLANS):
    """Launch shapes that minimise the residual time of an nk k-point gas on
    an ng-point grid: every power of two up to ng (and ng) as the chunk,
    then every GEMM plan at the best chunk, each timed over ``reps``
    residuals after one warm-up.  Returns (chunk, plan, {"chunk": {nb: s},
    "plan": {plan: s}})."""
    sysg = ueg_from_config(nk, rs, t_over_tf=t_over_tf)
    en = sysg.g_energies_tot()
    n = en.shape[0]
    beta = 1.0 / sysg.T
    ti, g, G = quadrature.ft_quad(ng, beta, 'lin')
    eng = FTEngine(n, n, ng, precision=precision)
    fock, eri, _ = sysg.g_device_integrals(eng)
    eng.set_system(en, fock, eri, beta, sysg.mu)
    del fock, eri
    eng.set_grid(ti, G)
    T1, T2 = eng.empty_amplitudes()
    T1b, T2b = eng.empty_amplitudes()
    eng.driving_terms(T1, T2)
    eng.integrate(T1, T2)
    eng.set_trusted(True)

    def timed():
        eng.residual(T1, T2, T1b, T2b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            eng.residual(T1, T2, T1b, T2b)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    cands = sorted({1 << k for k in range(ng.bit_length()) if 1 << k <= ng}
                   | {ng})
    chunk_times = {}
    for nb in cands:
        eng.set_chunk(nb)
        chunk_times[nb] = timed()
    best = min(chunk_times, key=chunk_times.get)
    eng.set_chunk(best)
    plan_times = {}
    for plan in plans:
        eng.set_gemm_plan(*plan)
        plan_times[plan] = timed()
    eng.close()
    plan = min(plan_times, key=plan_times.get)
    return best, plan, {"chunk": chunk_times, "plan": plan_times}


def autotune_chunk(nk, ng, precision="fp64", reps=3, rs=2.0, t_over_tf=0.5):
    """The chunk part of ``autotune`` alone: (best, {nb: seconds})."""
    best, _, times = autotune(nk, ng, precision, reps, rs, t_over_tf,
                              plans=((0, 1024),))
    return best, times["chunk"]


def cost(inst):
    """Relative cost of one iteration of an instance: n^6 ng."""
    return float(2 * inst["nk"]) ** 6 * inst["ng"]


def _solve_one(inst, precision, graph, chunk, plan, ccsd_kw, t_batch=None):
    sysg = ueg_from_config(inst["nk"], inst["rs"], N=inst["N"],
                           t_over_tf=inst["t_over_tf"])
    t = time.perf_counter()
    cc = ccsd(sysg, T=sysg.T, mu=sysg.mu, ngrid=inst["ng"],
              precision=precision, graph=graph, chunk=chunk, gemm_plan=plan,
              **ccsd_kw)
    Etot, Ecc = cc.run()
    torch.cuda.current_stream().synchronize()   # this worker's stream only
    done = time.perf_counter()
    iters = len(cc.history)
    hist = list(cc.history)
    mp2 = cc.mp2
    cc.engine.close()
    del cc, sysg
    out = dict(Omega=Etot, Omega_cc=Ecc, seconds=done - t, iterations=iters,
               history=hist, mp2=mp2)
    if t_batch is not None:
        out["latency"] = done - t_batch
    return out


def _trim_cache(limit_gb=24.0):
    """Hand torch's cached, unused device blocks back once they exceed
    ``limit_gb``: the library allocates with hipMalloc beside torch, and
    instances of other sizes do not reuse each other's blocks.  (Releasing
    synchronises the device, so it is done rarely rather than after every
    instance.)"""
    idle = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
    if idle > limit_gb * 2 ** 30:
        torch.cuda.empty_cache()


def solve_loop(instances, precision="fp64", progress=None, **ccsd_kw):
    """The plain loop: one instance after another on the default stream, in
    the given order; each result's ``latency`` is its completion time from
    the start of the loop."""
    t = time.perf_counter()
    res = []
    for k, x in enumerate(instances):
        res.append(_solve_one(x, precision, False, None, None, ccsd_kw, t))
        _trim_cache()
        if progress:
            progress(k + 1, len(instances))
    return res, time.perf_counter() - t


def solve_batch(instances, precision="fp64", workers=4, chunks=None,
                plans=None, **ccsd_kw):
    """Concurrent solves: ``workers`` (at most 4) threads, each with its own
    stream, every step a HIP graph, instances started shortest first;
    ``chunks`` and ``plans`` map (nk, ng) to a chunk size and a GEMM plan
    (see ``autotune``); the chunk is capped by each instance's share of HBM
    (0.35 of the device's memory over ``workers`` for the residual
    workspace, through FTCC_WORKSPACE_GB; amplitudes, integrals and torch's
    cache take the rest).  Returns (results in input order, wall
    seconds); each result's ``seconds`` is that instance's own solve time
    and ``latency`` its completion time from the start of the batch."""
    if not 1 <= workers <= 4:
        raise ValueError("solve_batch uses 1 to 4 worker streams (the "
                         "device's hardware queues per process)")
    chunks = chunks or {}
    plans = plans or {}
    total = torch.cuda.get_device_properties(0).total_memory
    saved = os.environ.get("FTCC_WORKSPACE_GB")
    os.environ["FTCC_WORKSPACE_GB"] = "%.3f" % (0.35 * total / workers /
                                                2 ** 30)
    order = sorted(range(len(instances)), key=lambda k: cost(instances[k]))
    t = time.perf_counter()

    def work(k):
        inst = instances[k]
        key = (inst["nk"], inst["ng"])
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            r = _solve_one(inst, precision, True, chunks.get(key),
                           plans.get(key), ccsd_kw, t)
        stream.synchronize()
        _trim_cache()
        return k, r

    res = [None] * len(instances)
    try:
        with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
            for k, r in ex.map(work, order):
                res[k] = r
    finally:
        if saved is None:
            os.environ.pop("FTCC_WORKSPACE_GB", None)
        else:
            os.environ["FTCC_WORKSPACE_GB"] = saved
    return res, time.perf_counter() - t


def _pct(x, q):
    return float(numpy.percentile(numpy.asarray(x), q))


def _stats(res, key):
    x = [r[key] for r in res]
    return {"p50_s": _pct(x, 50), "p99_s": _pct(x, 99)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--precision", default="fp64",
                    choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--damp", type=float, default=0.5)
    ap.add_argument("--max-iter", type=int, default=60)
    ap.add_argument("--econv", type=float, default=1e-8)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-loop", action="store_true",
                    help="skip the plain loop (no comparison)")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("the batch workload needs the GPU", file=sys.stderr)
        return 1
    insts = c4_instances(args.count, args.seed)
    kw = dict(damp=args.damp, max_iter=args.max_iter, econv=args.econv)
    chunks, plans, tune = {}, {}, {}
    t = time.perf_counter()
    if not args.no_autotune:
        for key in sorted({(x["nk"], x["ng"]) for x in insts}):
            best, plan, times = autotune(key[0], key[1], args.precision)
            chunks[key] = best
            plans[key] = plan
            tune["%d,%d" % key] = {
                "chunk": best, "plan": list(plan),
                "chunk_ms": {str(k): round(1e3 * v, 3)
                             for k, v in times["chunk"].items()},
                "plan_ms": {"%d/%d" % k: round(1e3 * v, 3)
                            for k, v in times["plan"].items()}}
            print("[batch] autotune nk=%d ng=%d: %s" % (
                key[0], key[1], json.dumps(tune["%d,%d" % key])),
                file=sys.stderr, flush=True)
    t_tune = time.perf_counter() - t
    # warm the library and allocator once
    _solve_one(insts[0], args.precision, False, None, None, kw)
    out = {"workload": "%d FT-CCSD UEG instances, n in {38,54,66}, ng in "
                       "{8,12,16,24}, T/T_F in [0.05,1], rs in [1,5], N=14 "
                       "(default_rng(%d))" % (args.count, args.seed),
           "precision": args.precision, "workers": args.workers,
           "autotune_s": round(t_tune, 2), "autotune": tune}
    if not args.no_loop:
        def prog(k, n):
            if k % 8 == 0 or k == n:
                print("[batch] loop %d/%d" % (k, n), file=sys.stderr,
                      flush=True)
        loop, t_loop = solve_loop(insts, args.precision, progress=prog, **kw)
        print("[batch] plain loop: %.2f s" % t_loop, file=sys.stderr,
              flush=True)
        out["loop"] = {"instances_per_s": args.count / t_loop,
                       "wall_s": t_loop, "solve": _stats(loop, "seconds"),
                       "latency": _stats(loop, "latency")}
    bat, t_bat = solve_batch(insts, args.precision, args.workers, chunks,
                             plans, **kw)
    print("[batch] batched: %.2f s" % t_bat, file=sys.stderr, flush=True)
    out["batched"] = {"instances_per_s": args.count / t_bat, "wall_s": t_bat,
                      "solve": _stats(bat, "seconds"),
                      "latency": _stats(bat, "latency")}
    out["iterations"] = [r["iterations"] for r in bat]
    if not args.no_loop:
        out["max_abs_dOmega_batched_vs_loop"] = max(
            abs(a["Omega"] - b["Omega"]) for a, b in zip(loop, bat))
        out["speedup"] = t_loop / t_bat
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
