"""Many FT-CCSD instances on one GPU: the serving-style workload.

A batch is a list of (T, mu, system) instances of uneven size n and grid
ng, each an ordinary ``ccsd(...).run()`` (the reference solves them one
after another; kelvin/ccsd.py has no batching).  ``solve_batch`` runs them
concurrently: each worker thread owns a HIP stream, and each instance runs
there with every fixed-point step replayed as a HIP graph
(``ccsd(graph=True)``), so the GPU overlaps the small instances' kernels and
the host issues one graph launch per iteration instead of a few hundred
kernel launches.  ``autotune_chunk`` picks, per (n, ng, precision), the
number of grid points whose products go out as one launch by timing the
residual at each choice; ``solve_batch`` applies it.

``python -m kelvin_amd.batch`` runs the 64-instance mix (n in {38, 54, 66},
i.e. 19, 27 or 33 k-points, ng in {8, 12, 16, 24}, T/T_F in [0.05, 1],
rs in [1, 5], N = 14, drawn from ``numpy.random.default_rng(4)``) both as
the plain loop and batched, and prints one JSON line with aggregate
instances/s, p50/p99 per-instance solve times and the largest difference
in Omega between the two.
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


def autotune_chunk(nk, ng, precision="fp64", reps=3, rs=2.0, t_over_tf=0.5):
    """Grid points per residual chunk that minimise the residual time of an
    nk k-point gas on an ng-point grid: every power of two up to ng (and
    ng) is timed over ``reps`` residuals after one warm-up.  Returns
    (best, {nb: seconds})."""
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
    cands = sorted({1 << k for k in range(ng.bit_length()) if 1 << k <= ng}
                   | {ng})
    times = {}
    for nb in cands:
        eng.set_chunk(nb)
        eng.residual(T1, T2, T1b, T2b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            eng.residual(T1, T2, T1b, T2b)
        torch.cuda.synchronize()
        times[nb] = (time.perf_counter() - t) / reps
    eng.close()
    best = min(times, key=times.get)
    return best, times


def _solve_one(inst, precision, graph, chunk, ccsd_kw):
    sysg = ueg_from_config(inst["nk"], inst["rs"], N=inst["N"],
                           t_over_tf=inst["t_over_tf"])
    t = time.perf_counter()
    cc = ccsd(sysg, T=sysg.T, mu=sysg.mu, ngrid=inst["ng"],
              precision=precision, graph=graph, chunk=chunk, **ccsd_kw)
    Etot, Ecc = cc.run()
    torch.cuda.current_stream().synchronize()   # this worker's stream only
    dt = time.perf_counter() - t
    iters = len(cc.history)
    cc.engine.close()
    # the library allocates with hipMalloc: hand torch's cached blocks of
    # this instance back before the next one
    del cc, sysg
    torch.cuda.empty_cache()
    return dict(Omega=Etot, Omega_cc=Ecc, seconds=dt, iterations=iters)


def solve_loop(instances, precision="fp64", progress=None, **ccsd_kw):
    """The plain loop: one instance after another on the default stream."""
    t = time.perf_counter()
    res = []
    for k, x in enumerate(instances):
        res.append(_solve_one(x, precision, False, None, ccsd_kw))
        if progress:
            progress(k + 1, len(instances))
    return res, time.perf_counter() - t


def solve_batch(instances, precision="fp64", workers=8, chunks=None,
                **ccsd_kw):
    """Concurrent solves: ``workers`` threads, each with its own stream,
    every step a HIP graph; ``chunks`` maps (nk, ng) to a chunk size (see
    ``autotune_chunk``), capped by each instance's share of HBM (half the
    device's memory over ``workers`` for the residual workspace, through
    FTCC_WORKSPACE_GB).  Returns (results in input order, wall seconds);
    each result's ``seconds`` is that instance's own solve time."""
    chunks = chunks or {}
    total = torch.cuda.get_device_properties(0).total_memory
    saved = os.environ.get("FTCC_WORKSPACE_GB")
    os.environ["FTCC_WORKSPACE_GB"] = "%.3f" % (0.5 * total / workers / 2 ** 30)

    def work(inst):
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            r = _solve_one(inst, precision, True,
                           chunks.get((inst["nk"], inst["ng"])), ccsd_kw)
        stream.synchronize()
        return r

    t = time.perf_counter()
    try:
        with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
            res = list(ex.map(work, instances))
    finally:
        if saved is None:
            os.environ.pop("FTCC_WORKSPACE_GB", None)
        else:
            os.environ["FTCC_WORKSPACE_GB"] = saved
    return res, time.perf_counter() - t


def _pct(x, q):
    return float(numpy.percentile(numpy.asarray(x), q))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--workers", type=int, default=8)
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
    chunks, tune = {}, {}
    t = time.perf_counter()
    if not args.no_autotune:
        for key in sorted({(x["nk"], x["ng"]) for x in insts}):
            best, times = autotune_chunk(key[0], key[1], args.precision)
            chunks[key] = best
            tune["%d,%d" % key] = {"best": best,
                                   "ms": {str(k): round(1e3 * v, 3)
                                          for k, v in times.items()}}
            print("[batch] autotune nk=%d ng=%d: chunk %d %s" % (
                key[0], key[1], best, tune["%d,%d" % key]["ms"]),
                file=sys.stderr, flush=True)
    t_tune = time.perf_counter() - t
    # warm the library and allocator once
    _solve_one(insts[0], args.precision, False, None, kw)
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
                       "wall_s": t_loop,
                       "p50_s": _pct([r["seconds"] for r in loop], 50),
                       "p99_s": _pct([r["seconds"] for r in loop], 99)}
    bat, t_bat = solve_batch(insts, args.precision, args.workers, chunks,
                             **kw)
    print("[batch] batched: %.2f s" % t_bat, file=sys.stderr, flush=True)
    out["batched"] = {"instances_per_s": args.count / t_bat, "wall_s": t_bat,
                      "p50_s": _pct([r["seconds"] for r in bat], 50),
                      "p99_s": _pct([r["seconds"] for r in bat], 99)}
    out["iterations"] = [r["iterations"] for r in bat]
    if not args.no_loop:
        out["max_abs_dOmega_batched_vs_loop"] = max(
            abs(a["Omega"] - b["Omega"]) for a, b in zip(loop, bat))
        out["speedup"] = t_loop / t_bat
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
