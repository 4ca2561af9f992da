"""Fermi-weighted integrals and the damped amplitude iteration.

Mirrors kelvin/cc_utils.py: ``ft_integrals`` (:569-602),
``ft_active_integrals`` (:825-858), ``form_new_ampl`` (:17-49) and
``ft_cc_iter`` (:111-173).  The weighted blocks are built on the GPU and kept
there; each iteration is one ``ftcc_step`` (residual at every grid point,
exp(-tau D) integration, damping, norms and energy fused on the device);
the host only keeps the loop's scalars.
"""
import logging
import math
import time

import numpy
import torch

from . import ft_cc_equations
from .engine import FTEngine
from .ov_blocks import engine_eri, engine_fock


def _engine_for_system(sys, en, beta, mu, iocc, ivir, ng, precision):
    n = len(en)
    no = n if iocc is None else len(iocc)
    nv = n if ivir is None else len(ivir)
    eng = FTEngine(no, nv, ng, precision=precision)
    eng.omega0 = eng.set_system(en, sys.g_fock_tot(), sys.g_aint_tot(), beta,
                                mu, iocc=iocc, ivir=ivir)
    return eng


def ft_integrals(sys, en, beta, mu, ng=1, precision="fp64"):
    """Return (F, I): handles on the GPU-resident sqrt(f)-weighted Fock and
    ERI blocks over all n orbitals (kelvin/cc_utils.py:569-602)."""
    eng = _engine_for_system(sys, en, beta, mu, None, None, ng, precision)
    return engine_fock(eng), engine_eri(eng)


def ft_active_integrals(sys, en, focc, fvir, iocc, ivir, beta=None, mu=None,
                        ng=1, precision="fp64"):
    """Blocks restricted to the orbitals with f > athresh
    (kelvin/cc_utils.py:825-858).  The device recomputes the Fermi factors
    from (beta, mu); focc/fvir are accepted for signature compatibility."""
    if beta is None or mu is None:
        T, mu = getattr(sys, "T"), getattr(sys, "mu")
        beta = 1.0 / T
    eng = _engine_for_system(sys, en, beta, mu, list(iocc), list(ivir), ng,
                             precision)
    return engine_fock(eng), engine_eri(eng)


def form_new_ampl(method, F, I, T1old, T2old, D1, D2, ti, ng, G):
    """New (integrated) amplitudes (kelvin/cc_utils.py:17-49)."""
    if method == "CCSD":
        return ft_cc_equations.ccsd_stanton(F, I, T1old, T2old, D1, D2, ti,
                                            ng, G)
    if method in ("CCD", "LCCSD", "LCCD"):
        raise NotImplementedError(
            "method %s is not built for the GPU path" % method)
    raise Exception("Unrecognized method keyword")


def zero_slices(engine, T1, T2, G):
    """Bit mask of grid points whose amplitudes are exactly zero and stay so
    (G[y,:] == 0): their residual is the driving term alone."""
    mask = 0
    Gm = numpy.asarray(G)
    for y in range(min(T1.shape[0], 64)):
        if numpy.all(Gm[y] == 0.0):
            if engine.sumsq(T1[y]) == 0.0 and engine.sumsq(T2[y]) == 0.0:
                mask |= 1 << y
    return mask


def ft_cc_iter(method, T1old, T2old, F, I, D1, D2, g, G, beta, ng, ti,
               iprint, conv_options, history=None):
    """Damped fixed-point iteration (kelvin/cc_utils.py:111-173).

    T1old/T2old are device tensors updated in place; returns
    (Eold, T1old, T2old).  ``history`` (optional list) receives
    (E, res) per iteration.
    """
    if method != "CCSD":
        return _other_method(method)
    eng = ft_cc_equations.engine_of(F, I, D1, D2, ng)
    eng.set_grid(ti, G)
    tbeg = time.time()
    converged = False
    ethresh = conv_options["econv"]
    tthresh = conv_options["tconv"]
    max_iter = conv_options["max_iter"]
    alpha = conv_options["damp"]
    T1b = torch.empty_like(T1old)
    T2b = torch.empty_like(T2old)
    skip = zero_slices(eng, T1old, T2old, G)
    i = 0
    Eold = 888888888.888888888
    nl1 = math.sqrt(eng.sumsq(T1old)) + 0.1
    nl2 = math.sqrt(eng.sumsq(T2old)) + 0.1
    graph = getattr(eng, "use_graph", False)
    if graph:
        eng.set_weights(g)
    while i < max_iter and not converged:
        if graph:
            d1, d2, n1, n2, E = eng.step_graph(T1old, T2old, T1b, T2b, beta,
                                               alpha, skip)
        else:
            d1, d2, n1, n2, E = eng.step(T1old, T2old, T1b, T2b, g, beta,
                                         alpha, skip)
        res1 = math.sqrt(d1) / nl1
        res2 = math.sqrt(d2) / nl2
        nl1 = math.sqrt(n1) + 0.1
        nl2 = math.sqrt(n2) + 0.1
        logging.info(' %2d  %.10f   %.4E' % (i + 1, E, res1 + res2))
        if history is not None:
            history.append((E, res1 + res2))
        i = i + 1
        if abs(E - Eold) < ethresh and res1 + res2 < tthresh:
            converged = True
        Eold = E
    if not converged:
        logging.warning("{} did not converge!".format(method))
    tend = time.time()
    logging.info("Total {} time: {:.4f} s".format(method, (tend - tbeg)))
    del T1b, T2b
    return Eold, T1old, T2old


def ft_cc_iter_refined(method, T1old, T2old, F, I, D1, D2, g, G, beta, ng,
                       ti, iprint, conv_options, low, history=None,
                       stats=None):
    """Mixed-precision damped fixed point with an FP64 finish.

    The loop, norms, damping and convergence test are those of
    ``ft_cc_iter`` (kelvin/cc_utils.py:111-173) in three phases:

    1. residuals in the low precision ``low`` ('fp32' or 'bf16x3') until the
       relative change falls below 10 tconv or, once below 1e-4, stops
       decreasing (rounding of the low-precision products sets a floor
       under it);
    2. the defect dTbar = Tbar_fp64(T_c) - Tbar_low(T_c) at the current
       iterate T_c is added to every further low-precision residual, which
       moves the fixed point from the low-precision one (about 1e-7 Ha away
       in bf16x3 at n = 114) to the FP64 one, until the change falls below
       tconv or stops decreasing again;
    3. plain FP64 steps until the reference's test |E - Eold| < econv and
       res < tconv holds, so that convergence is always certified by an FP64
       step.

    ``stats`` (a dict) receives the number of steps in each phase.  T1old /
    T2old are updated in place; returns (Eold, T1old, T2old).
    """
    if method != "CCSD":
        return _other_method(method)
    eng = ft_cc_equations.engine_of(F, I, D1, D2, ng)
    eng.set_grid(ti, G)
    eng.set_precision(low)
    tbeg = time.time()
    ethresh = conv_options["econv"]
    tthresh = conv_options["tconv"]
    max_iter = conv_options["max_iter"]
    alpha = conv_options["damp"]
    T1b = torch.empty_like(T1old)
    T2b = torch.empty_like(T2old)
    dT1 = dT2 = None
    skip = zero_slices(eng, T1old, T2old, G)
    graph = getattr(eng, "use_graph", False)
    if graph:
        eng.set_weights(g)
    nl1 = math.sqrt(eng.sumsq(T1old)) + 0.1
    nl2 = math.sqrt(eng.sumsq(T2old)) + 0.1
    Eold = 888888888.888888888
    i, phase = 0, 1
    steps = {1: 0, 2: 0, 3: 0}
    best, flat = float("inf"), 0
    converged = False
    try:
        while i < max_iter and not converged:
            if graph:
                d1, d2, n1, n2, E = eng.step_graph(T1old, T2old, T1b, T2b,
                                                   beta, alpha, skip)
            else:
                d1, d2, n1, n2, E = eng.step(T1old, T2old, T1b, T2b, g, beta,
                                             alpha, skip)
            res = math.sqrt(d1) / nl1 + math.sqrt(d2) / nl2
            nl1 = math.sqrt(n1) + 0.1
            nl2 = math.sqrt(n2) + 0.1
            logging.info(' %2d  %.10f   %.4E' % (i + 1, E, res))
            if history is not None:
                history.append((E, res))
            i += 1
            steps[phase] += 1
            ok = abs(E - Eold) < ethresh and res < tthresh
            Eold = E
            if phase == 3:
                converged = ok
                continue
            if res < 0.9 * best:
                best, flat = res, 0
            else:
                flat += 1
            near = res < 1e-4        # past the transient of the iteration
            if phase == 1 and (ok or res < 10.0 * tthresh or
                               (near and flat >= 5)):
                t = time.time()
                dT1 = torch.empty_like(T1old)
                dT2 = torch.empty_like(T2old)
                eng.defect(T1old, T2old, low, dT1, dT2, T1b, T2b, skip)
                eng.set_correction(dT1, dT2)
                phase, best, flat = 2, float("inf"), 0
                logging.info("FP64 defect at step %d: %.3f s" %
                             (i, time.time() - t))
            elif phase == 2 and (ok or res < tthresh or
                                 (flat >= 4 and steps[2] >= 6)):
                eng.set_correction(None, None)
                del dT1, dT2
                dT1 = dT2 = None
                eng.set_precision("fp64")
                phase = 3
                logging.info("FP64 finish from step %d" % i)
    finally:
        eng.set_correction(None, None)
    if not converged:
        logging.warning("{} did not converge!".format(method))
    logging.info("Total {} time: {:.4f} s ({} {} steps, {} corrected, {} "
                 "FP64)".format(method, time.time() - tbeg, steps[1], low,
                                steps[2], steps[3]))
    if stats is not None:
        stats["steps"] = dict(steps)
        stats["converged"] = converged
    del T1b, T2b, dT1, dT2
    return Eold, T1old, T2old


def ft_cc_iter_extrap(method, F, I, D1, D2, g, G, beta, ng, ti,
                      iprint, conv_options):
    """Pointwise schedule (kelvin/cc_utils.py:176-242): converge the grid
    points one at a time, each started from the integrated driving terms
    (point 1) or a linear extrapolation of the previous two points, with
    damping and the RMS-change test res1 + res2 < tconv.  Returns device
    tensors (T1, T2)."""
    if method != "CCSD":
        return _other_method(method)
    eng = ft_cc_equations.engine_of(F, I, D1, D2, ng)
    eng.set_grid(ti, G)
    thresh = conv_options["tconv"]
    max_iter = conv_options["max_iter"]
    alpha = conv_options["damp"]
    T1, T2 = eng.empty_amplitudes()
    T1.zero_()
    T2.zero_()
    T1b = torch.zeros_like(T1)
    T2b = torch.zeros_like(T2)
    nl1 = math.sqrt(float(eng.nv * eng.no))
    nl2 = math.sqrt(float(eng.nv * eng.nv * eng.no * eng.no))
    for ig in range(ng):
        eng.point_begin(ig, T1, T2, T1b, T2b)
        if ig == 0:
            continue
        logging.info("Time point {}".format(ig))
        i, converged = 0, False
        while i < max_iter and not converged:
            d1, d2 = eng.point_step(ig, T1, T2, T1b, T2b, alpha)
            res = math.sqrt(d1) / nl1 + math.sqrt(d2) / nl2
            logging.info(' %2d  %.4E' % (i + 1, res))
            i += 1
            converged = res < thresh
    del T1b, T2b
    return T1, T2


def uft_integrals(sys, ea, eb, beta, mu):
    """Unrestricted Fermi-weighted blocks (kelvin/cc_utils.py:696-777):
    (Fa, Fb, Ia, Ib, Iabab) as device tensors.  The spin-orbital Fock matrix
    and antisymmetrised integrals are assembled from sys.u_fock_tot() and
    sys.u_aint_tot(), weighted on the device like the spin-orbital path, and
    the alpha/beta blocks are taken back out."""
    from . import unrestricted
    ea = numpy.asarray(ea, dtype=numpy.float64)
    eb = numpy.asarray(eb, dtype=numpy.float64)
    na, nb = ea.shape[0], eb.shape[0]
    n = na + nb
    fa, fb = sys.u_fock_tot()
    eA, eB, eAB = sys.u_aint_tot()
    fock = numpy.zeros((n, n))
    fock[:na, :na] = fa
    fock[na:, na:] = fb
    eri = numpy.zeros((n, n, n, n))
    a, b = slice(0, na), slice(na, n)
    eri[a, a, a, a] = eA
    eri[b, b, b, b] = eB
    eri[a, b, a, b] = eAB
    eri[a, b, b, a] = -eAB.transpose(0, 1, 3, 2)
    eri[b, a, a, b] = -eAB.transpose(1, 0, 2, 3)
    eri[b, a, b, a] = eAB.transpose(1, 0, 3, 2)
    eng = FTEngine(n, n, 1)
    eng.set_system(numpy.concatenate([ea, eb]), fock, eri, beta, mu)
    del eri
    F, I = engine_fock(eng), engine_eri(eng)
    lay = unrestricted.SpinLayout(na, nb, na, nb)
    return unrestricted.unrestricted_blocks(lay, F, I)


def ft_ucc_iter(method, T1aold, T1bold, T2aaold, T2abold, T2bbold, Fa, Fb,
                Ia, Ib, Iabab, D1a, D1b, D2aa, D2ab, D2bb, g, G, beta, ng, ti,
                iprint, conv_options):
    """Unrestricted damped fixed point (kelvin/cc_utils.py:245-317), with
    the reference's spin-block norms: nl1 = |T1a| + 0.1 + |T1b|,
    nl2 = |T2aa| + 0.1 + |T2ab| + |T2bb|, res = sum of block changes over
    those.  Returns (E, (T1a, T1b), (T2aa, T2ab, T2bb))."""
    from . import unrestricted
    if method != "CCSD":
        return _other_method(method)
    eng, lay = ft_cc_equations.unrestricted_engine(
        Fa, Fb, Ia, Ib, Iabab, D1a, D1b, D2aa, D2ab, D2bb, ng)
    eng.set_grid(ti, G)
    T1, T2 = unrestricted.spin_orbital_amplitudes(
        lay, T1aold, T1bold, T2aaold, T2abold, T2bbold, eng.device)
    T1n = torch.empty_like(T1)
    T2n = torch.empty_like(T2)
    skip = zero_slices(eng, T1, T2, G)
    ethresh = conv_options["econv"]
    tthresh = conv_options["tconv"]
    max_iter = conv_options["max_iter"]
    alpha = conv_options["damp"]

    def nrm(x):
        return math.sqrt(eng.sumsq(x.contiguous()))

    tbeg = time.time()
    converged, i, Eold = False, 0, 888888888.888888888
    while i < max_iter and not converged:
        eng.residual(T1, T2, T1n, T2n, skip)
        eng.integrate(T1n, T2n)
        o1, o2 = unrestricted.unrestricted_amplitudes(lay, T1, T2)
        n1, n2 = unrestricted.unrestricted_amplitudes(lay, T1n, T2n)
        nl1 = nrm(o1[0]) + 0.1 + nrm(o1[1])
        nl2 = nrm(o2[0]) + 0.1 + nrm(o2[1]) + nrm(o2[2])
        res1 = sum(nrm(x - y) for x, y in zip(n1, o1)) / nl1
        res2 = sum(nrm(x - y) for x, y in zip(n2, o2)) / nl2
        T1.mul_(alpha).add_(T1n, alpha=1.0 - alpha)
        T2.mul_(alpha).add_(T2n, alpha=1.0 - alpha)
        E = eng.energy(T1, T2, g, beta, qterm=True)
        logging.info(' %2d  %.10f   %.4E' % (i + 1, E, res1 + res2))
        i += 1
        if abs(E - Eold) < ethresh and res1 + res2 < tthresh:
            converged = True
        Eold = E
    if not converged:
        logging.warning("{} did not converge!".format(method))
    logging.info("Total {} time: {:.4f} s".format(method, time.time() - tbeg))
    T1u, T2u = unrestricted.unrestricted_amplitudes(lay, T1, T2)
    return Eold, T1u, T2u


def _other_method(method):
    if method in ("CCD", "LCCSD", "LCCD"):
        raise NotImplementedError(
            "method %s is not built for the GPU path" % method)
    raise Exception("Unrecognized method keyword")
