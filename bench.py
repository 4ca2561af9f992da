#!/usr/bin/env python3
"""FT-CCSD amplitude-iteration benchmark on one MI355X.

Workload: uniform electron gas, 57 k-points (n = 114 spin orbitals), rs = 2,
T = 0.1 T_F, N = 14 (mu solved for <N> = 14), ng = 24 Simpson grid.  One step
is one body of the damped fixed point (residual at every grid point,
exp(-tau D) integration, damping, norms, grand potential), with amplitudes
resident in HBM.  The headline value is mixed-precision (default bf16x3)
iterations/s; the FP64 form is timed from the same amplitudes and the
grand-potential difference of that step is reported.

  python bench.py [--steps K] [--warmup W] [--precision bf16x3|fp32|fp64]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("FT-CCSD iters/sec at n=114, ng=24: mixed vs FP64 GPU; "
          "|ΔΩ| vs FP64 oracle")
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
MFMA_PER_PRODUCT = {"fp64": 1, "fp32": 1, "bf16x3": 3}


def peak_of(args):
    return PEAK_TFLOPS[args.precision]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="bf16x3",
                    choices=["bf16x3", "fp32", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--fp32-steps", type=int, default=1)
    ap.add_argument("--chunk", type=int, default=0,
                    help="grid points per residual chunk (0: the library's "
                         "choice from its workspace budget)")
    ap.add_argument("--gemm-plan", default=None,
                    help="tile_policy/split_target of the generic GEMMs "
                         "(default: the library's)")
    ap.add_argument("--nk", type=int, default=57)
    ap.add_argument("--ng", type=int, default=24)
    ap.add_argument("--rs", type=float, default=2.0)
    ap.add_argument("--t-over-tf", type=float, default=0.1)
    ap.add_argument("--nelec", type=int, default=14)
    ap.add_argument("--damp", type=float, default=0.3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=90.0,
                    help="predicted CPU sample time above which the sample "
                         "is taken at a smaller n and extrapolated")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the warmup, copy a buffer of this many MiB "
                         "once on the device: the known-size copy kernel that "
                         "hardware byte counters are scaled by")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sbgemm_traffic_ueg_n114.txt")


KERNEL_SOURCE = os.path.join(ROOT, "kelvin_amd", "csrc", "sbgemm.hip")


def source_digest(path):
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def measured_traffic(precision, n, ng, layout):
    """HBM bytes per launch of the dominant kernel, from the committed
    FETCH_SIZE/WRITE_SIZE profile of this workload (a line
    'traffic_bytes_per_launch <precision> <n> <ng> <layout> <bytes>',
    layout e.g. 'spin-packed-flip-c8': the operand layout, '-flip' when the
    ring products compute the alpha-virtual rows only, and the grid points
    per launch), or None.  The profile names the digest of the kernel
    source it was taken with ('kernel_source_sha256 <16 hex>'); a profile
    of another source is not used: hardware counters cannot be read from
    inside this process, so a stale figure is dropped, not reported."""
    try:
        digest = source_digest(KERNEL_SOURCE)
        found, same_source = None, False
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 2 and w[0] == "kernel_source_sha256":
                    same_source = w[1] == digest
                if (len(w) == 6 and w[0] == "traffic_bytes_per_launch"
                        and w[1] == precision and int(w[2]) == n
                        and int(w[3]) == ng and w[4] == layout):
                    found = float(w[5])
        return found if same_source else None
    except OSError:
        pass
    return None


def host_description():
    """CPU model, logical CPUs and the BLAS numpy calls (the cpu leg's
    hardware and library)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = None
    try:
        from threadpoolctl import threadpool_info
        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            blas = "%s %s (%s, %d threads)" % (
                info[0].get("internal_api"), info[0].get("version"),
                info[0].get("threading_layer"), info[0]["num_threads"])
    except Exception:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu": model, "logical_cpus": os.cpu_count(),
            "cpus_allowed": affinity, "blas": blas}


def cpu_threads():
    try:
        from threadpoolctl import threadpool_info
        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            return int(info[0]["num_threads"])
    except Exception:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(eng, T1, T2, n, ng, budget_s):
    """Time the oracle's residual (numpy FP64, host cores) for one grid point
    at the benchmark's own n and scale to one iteration (ng grid points)."""
    import numpy
    import torch
    from oracle import ftcc as oracle
    from kelvin_amd.engine import ONE_E, TWO_E
    # rate probe: one 1/16 row slab of an n^2 x n^2 x n^2 product
    m = n * n
    a = numpy.random.default_rng(0).normal(size=(max(1, m // 16), m))
    b = numpy.random.default_rng(1).normal(size=(m, m))
    t = time.perf_counter()
    a @ b
    probe = time.perf_counter() - t
    rate = 2.0 * a.shape[0] * m * m / probe
    del a, b
    predicted = 12.0 * n ** 6 / rate * 1.3
    log("cpu rate probe %.1f GFLOP/s, predicted one-point residual %.1f s"
        % (rate / 1e9, predicted))
    if predicted <= budget_s:
        F = oracle.Blocks({k: eng.get_block(k).cpu().numpy() for k in ONE_E})
        I = oracle.Blocks({k: eng.get_block(k).cpu().numpy() for k in TWO_E})
        t1 = T1[ng - 1].cpu().numpy()
        t2 = T2[ng - 1].cpu().numpy()
        t = time.perf_counter()
        oracle.stanton_rhs(F, I, t1, t2)
        dt = time.perf_counter() - t
        sample = ("oracle stanton_rhs (numpy FP64) at n=%d for one grid point, "
                  "%.1f s, scaled x%d grid points per iteration" % (n, dt, ng))
        return 1.0 / (dt * ng), sample
    # too slow for the budget: time one grid point at a smaller gas and
    # extrapolate by the n^6 cost of the residual
    from oracle import ftcc as o
    nk_s = 19
    L, T, Ef = o.ueg_params(nk_s, 2.0, 14, 0.1)
    s = o.UEGSystem(T, L, 0.5 * (2 * math.pi / L) ** 2 * 8.0,
                    0.2, norb=nk_s)
    en = s.g_energies_tot()
    F, I = o.ft_integrals(s.g_fock_tot(), s.g_aint_tot(), en, 1.0 / T, 0.2)
    ns = en.shape[0]
    rng = numpy.random.default_rng(2)
    t1 = rng.normal(size=(ns, ns)) * 1e-3
    t2 = rng.normal(size=(ns, ns, ns, ns)) * 1e-3
    t = time.perf_counter()
    o.stanton_rhs(F, I, t1, t2)
    dt = time.perf_counter() - t
    dt_full = dt * (float(n) / ns) ** 6
    sample = ("EXTRAPOLATED: oracle stanton_rhs at n=%d for one grid point "
              "took %.2f s; scaled by (%d/%d)^6 and x%d grid points"
              % (ns, dt, n, ns, ng))
    return 1.0 / (dt_full * ng), sample


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU "
              "path); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy
    import torch
    from kelvin_amd import quadrature
    from kelvin_amd.cc_utils import zero_slices
    from kelvin_amd.engine import FTEngine
    from kelvin_amd.ueg_system import ueg_from_config

    if not torch.cuda.is_available():
        print("bench.py needs the GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    t0 = time.time()
    sysg = ueg_from_config(args.nk, args.rs, N=args.nelec,
                           t_over_tf=args.t_over_tf)
    en = sysg.g_energies_tot()
    n = en.shape[0]
    beta, mu = 1.0 / sysg.T, sysg.mu
    ng = args.ng
    ti, g, G = quadrature.ft_quad(ng, beta, 'lin')
    eng = FTEngine(n, n, ng, precision=args.precision)
    if args.chunk:
        eng.set_chunk(args.chunk)
    if args.gemm_plan:
        eng.set_gemm_plan(*[int(x) for x in args.gemm_plan.split("/")])
    fock, eri, e1 = sysg.g_device_integrals(eng)     # built in HBM
    log("UEG nk=%d n=%d beta=%.4f mu=%.6f, integrals built on the device "
        "in %.1f s" % (args.nk, n, beta, mu, time.time() - t0))
    omega0 = eng.set_system(en, fock, eri, beta, mu)
    eng.set_grid(ti, G)
    del eri, fock
    T1, T2 = eng.empty_amplitudes()
    T1b, T2b = eng.empty_amplitudes()
    eng.driving_terms(T1, T2)
    eng.integrate(T1, T2)
    skip = zero_slices(eng, T1, T2, G)
    # the iterates are the library's own (driving terms, integration,
    # steps): no per-step spin check of the amplitudes
    eng.set_trusted(True)
    spin_mode = eng.spin_mode()
    spin_blocked = bool(spin_mode & 1)
    # operand layout of the dominant kernel: dense, spin blocks, or spin
    # blocks with the ladder products over packed antisymmetric pairs
    layout = ("dense" if not spin_blocked else
              "spin-packed" if spin_mode & 4 else "spin")
    if spin_blocked and spin_mode & 8:
        layout += "-flip"    # ring products over the alpha-virtual rows
    log("device set-up done in %.1f s (context %.1f GB, product layout: %s, "
        "spin flags %d)" % (time.time() - t0, eng.workspace_bytes() / 1e9,
                            layout, spin_mode))
    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy of %d MiB done" % args.copy_calibration_mib)
        del src, dst

    for w in range(args.warmup):
        t = time.time()
        out = eng.step(T1, T2, T1b, T2b, g, beta, args.damp, skip)
        log("warmup %d: E=%.10f (%.2f s)" % (w + 1, out[4], time.time() - t))
    save1, save2 = T1.clone(), T2.clone()

    eng.profile(True)
    eng.profile_classes()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    energies = []
    for k in range(args.steps):
        out = eng.step(T1, T2, T1b, T2b, g, beta, args.damp, skip)
        energies.append(out[4])
        log("step %d: E=%.10f" % (k + 1, out[4]))
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t_start
    classes = eng.profile_classes()
    eng.profile(False)
    launches, gemm_ms, gemm_flops, _ = classes["products"]
    value = args.steps / t_mixed
    # the chunk of the timed steps (the other forms below plan their own)
    chunk = eng.chunk()

    def other_form(prec, steps):
        """The same step in another precision from the same amplitudes:
        one untimed step (its workspace and operand buffers are built on
        first use), then ``steps`` timed ones."""
        eng.set_precision(prec)
        # a solve in this precision alone plans its own chunk: release the
        # workspace (a chunk-setting round trip) before the untimed step
        cur = args.chunk or 0
        eng.set_chunk(cur + 1)
        eng.set_chunk(cur)
        T1.copy_(save1)
        T2.copy_(save2)
        eng.step(T1, T2, T1b, T2b, g, beta, args.damp, skip)
        T1.copy_(save1)
        T2.copy_(save2)
        torch.cuda.synchronize()
        t = time.perf_counter()
        es = []
        for k in range(steps):
            out = eng.step(T1, T2, T1b, T2b, g, beta, args.damp, skip)
            es.append(out[4])
            log("%s step %d: E=%.10f" % (prec, k + 1, out[4]))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        eng.set_precision(args.precision)
        return ({"value": steps / dt, "unit": "iter/s",
                 "ms_per_step": 1e3 * dt / steps}, es[0])

    fp64 = fp32 = None
    d_omega = d_omega32 = None
    e64 = None
    if args.precision != "fp64" and args.fp64_steps > 0:
        fp64, e64 = other_form("fp64", args.fp64_steps)
        d_omega = abs(energies[0] - e64)
    if args.precision != "fp32" and args.fp32_steps > 0:
        fp32, e32 = other_form("fp32", args.fp32_steps)
        if e64 is not None:
            d_omega32 = abs(e32 - e64)
    del save1, save2

    avg_ms = gemm_ms / max(1, launches)
    launch_flops = gemm_flops / max(1, launches)
    achieved = launch_flops / (avg_ms * 1e-3) / 1e12 if launches else 0.0
    peak = peak_of(args)
    roofline = {
        "bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
        "unit": "TFLOP/s", "frac": round(achieved / peak, 5),
        "traffic": measured_traffic(args.precision, n, ng,
                                    "%s-c%d" % (layout, chunk)),
        "kernel": "%s (the five O(n^6) products of the residual)"
                  % ("sbgemm_kernel" if args.precision == "bf16x3"
                     else "gemm_mfma_kernel<%s>" % ("double" if args.precision == "fp64" else "float")),
        "launch_flops": launch_flops, "avg_launch_ms": round(avg_ms, 4),
        "launches": launches,
        "mfma_ops_per_product": MFMA_PER_PRODUCT[args.precision],
    }
    # every kernel class of the step against the roofline that bounds it:
    # the O(n^5) GEMMs run in FP64 in every precision; an arithmetic
    # intensity above the FP64 ridge (78.6 TFLOP/s / 8 TB/s) makes a class
    # MFMA-bound, else HBM-bound
    per_class = {}
    for name, (nl, ms, fl, by) in classes.items():
        if nl == 0 or ms <= 0.0:
            continue
        sec = ms * 1e-3
        entry = {"launches_per_step": nl / args.steps,
                 "ms_per_step": ms / args.steps,
                 "avg_launch_ms": ms / nl}
        if name == "products":
            entry.update(bound="mfma", unit="TFLOP/s",
                         achieved=fl / sec / 1e12, peak=peak_of(args))
        elif fl > 0.0 and by > 0.0 and fl / by > 78.6 / 8.0:
            entry.update(bound="mfma", unit="TFLOP/s",
                         achieved=fl / sec / 1e12, peak=78.6,
                         hbm_gbs=by / sec / 1e9)
        else:
            entry.update(bound="hbm", unit="GB/s", achieved=by / sec / 1e9,
                         peak=8000.0)
            if fl > 0.0:
                entry["tflops"] = fl / sec / 1e12
        entry["frac"] = entry["achieved"] / entry["peak"]
        entry["algorithmic_flops_per_step"] = fl / args.steps
        entry["algorithmic_bytes_per_step"] = by / args.steps
        per_class[name] = entry
    cpu = None
    if not args.no_cpu_baseline:
        try:
            rate, sample = cpu_baseline(eng, T1, T2, n, ng, args.cpu_budget_s)
            cpu = {"value": rate, "unit": "iter/s", "cores": cpu_threads(),
                   "kind": "port", "sample": sample,
                   "host": host_description()}
        except Exception as exc:   # report, do not hide the GPU numbers
            cpu = {"value": None, "unit": "iter/s", "cores": cpu_threads(),
                   "kind": "port", "sample": "failed: %r" % (exc,),
                   "host": host_description()}
    algo = 12.0 * n ** 6 * ng
    line = {
        "metric": METRIC, "value": value, "unit": "iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (deterministic UEG integrals; mu solved for <N>)",
        "config": {"workload": "UEG %d k-points (n=%d spin orbitals), rs=%g, "
                               "T=%g T_F, N=%d, ng=%d Simpson grid"
                               % (args.nk, n, args.rs, args.t_over_tf,
                                  args.nelec, ng),
                   "n": n, "ng": ng, "precision": args.precision,
                   "damp": args.damp, "spin_blocked": spin_blocked,
                   "layout": layout, "chunk": chunk},
        "fp64": fp64,
        "fp32": fp32,
        "omega_step_delta_vs_fp64": d_omega,
        "omega_step_delta_fp32_vs_fp64": d_omega32,
        "textbook_tflops": algo * value / 1e12,
        "product_flops_per_iter": gemm_flops / args.steps,
        "omega0_plus_e1": omega0 + e1,
        "energies": energies,
        "roofline": roofline,
        "kernel_classes": per_class,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
