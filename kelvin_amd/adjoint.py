"""Adjoint of the converged FT-CCSD fixed point: lambda amplitudes, the
derivatives of the grand potential with beta and mu (E, S, N), and a
``torch.autograd.Function`` around them.

kelvin differentiates its discretised grand potential with lambda
amplitudes and response densities (``compute_ESN``, kelvin/ccsd.py:128-214;
``_ft_ccsd_lambda`` :882-961; ``ft_lambda_iter``, kelvin/cc_utils.py:
414-480; ``_g_ft_ron`` and ``_g_nocc_gderiv``, kelvin/ccsd.py:1075-1148,
1434-1492).  The same derivative is taken here as the adjoint of the fixed
point T = Phi(T; beta, mu) (Phi: residual at every grid point, then the
exp(-tau D) integration):

    dOmega_cc/dtheta = dE/dtheta + sum_y <Lam_y, dPhi_y/dtheta>,
    Lam = dE/dT + (dPhi/dT)^T Lam,

solved with kelvin's damped iteration and convergence test (the adjoint
amplitudes Lam are kelvin's L up to a factor per grid point:
Lam1_y = -(g_y/beta) L1_y^T, Lam2_y = -(g_y/(4 beta)) L2_y^T).  theta moves
the Fermi weights of the integrals and, through the occupations, the Fock
matrix (x_p = beta (e_p - mu)); beta also scales the grid.  The residual is
linear in the integrals, so the integral part is one residual evaluation
with the derivative blocks (ftcc_set_system_response), contracted with the
integrated adjoint; the grid part is the derivative of the weights and
propagators (ftcc_grid_derivative) and of the quadrature weights.  Every
step runs in libftcc on the GPU.
"""
import logging
import math
import time

import numpy
import torch

from . import ft_utils
from . import quadrature
from .cc_utils import zero_slices


def _norm_scaled(eng, L, w):
    """sqrt(sum_y w_y^2 |L_y|^2)."""
    return math.sqrt(sum(wy * wy * eng.sumsq(L[y]) for y, wy in enumerate(w)))


def solve_lambda(eng, T1, T2, g, beta, conv, L1=None, L2=None, history=None,
                 low=None, stats=None):
    """Adjoint amplitudes of the converged T (kelvin/cc_utils.py:414-480
    ft_lambda_iter with the guess of kelvin/ft_cc_equations.py:502-512, or
    kelvin's L1 / L2 as the start): L <- damp L + (1 - damp) L_new with
    L_new = dE/dT - J^T(Int^T L) at every grid point, norms |L| + 0.1 and
    the test res1 + res2 < tconv, all in kelvin's lambda scale.  The
    amplitude-only intermediates of J^T are built once (kept for as many
    grid points as FTCC_ADJOINT_CACHE_GB allows, rebuilt from T for the
    others).

    ``low`` ("bf16x3") runs the refined form.  J^T is linear in Lt, so
    (1) iterations with the large blocks of J^T as split-bf16 products run
    until the residual is below max(100 tconv, 1e-6) or stalls below 1e-3;
    (2) one FP64 application at that iterate, J^T_64(Lt0), becomes the
    anchor of the following steps, J^T(Lt) = J^T_64(Lt0) + J^T_low(Lt - Lt0),
    whose rounding error scales with |Lt - Lt0| rather than |Lt|; (3) once
    kelvin's test holds, an FP64 step decides it: converged when the test
    holds for that step too, otherwise that step is the new anchor.
    ``stats`` receives the steps of each kind (1 low, 2 anchored, 3 FP64).
    Returns (Lam1, Lam2, converged, iterations)."""
    g = numpy.asarray(g, dtype=numpy.float64)
    w1 = beta / g
    w2 = 4.0 * beta / g
    thresh, max_iter, damp = conv["tconv"], conv["max_iter"], conv["damp"]
    eng.adjoint_begin(T1, T2)
    eng.adjoint_slices_kept = eng.adjoint_cached()
    try:
        if L1 is None and L2 is None:
            Lam1, Lam2 = eng.empty_amplitudes()
            eng.lambda_guess(T1, g, beta, Lam1, Lam2)
        else:
            ng, nv, no = eng.ng, eng.nv, eng.no
            if L1 is None:
                L1 = torch.zeros((ng, no, nv), dtype=torch.float64)
            if L2 is None:
                L2 = torch.zeros((ng, no, no, nv, nv), dtype=torch.float64)
            Lam1, Lam2 = eng.lambda_from_kelvin(L1, L2, g, beta)
        nl1 = _norm_scaled(eng, Lam1, w1) + 0.1
        nl2 = _norm_scaled(eng, Lam2, w2) + 0.1
        Lt1, Lt2 = eng.empty_amplitudes()
        G1, G2 = eng.empty_amplitudes()
        i, converged = 0, False
        phase = 1 if low else 3
        steps = {1: 0, 2: 0, 3: 0}
        anchor = None     # (Lt1, Lt2, G1, G2) of the last FP64 application
        anchors = 0
        switch = max(100.0 * thresh, 1e-6)
        best, flat = float("inf"), 0
        if low:
            eng.adjoint_precision(low)
        tbeg = time.time()
        while i < max_iter and not converged:
            eng.integrate_adjoint(Lam1, Lam2, Lt1, Lt2)
            if phase == 3 or (phase == 2 and anchor is None):
                kind = 3
                if low:
                    eng.adjoint_precision("fp64")
                eng.adjoint_apply(Lt1, Lt2, G1, G2)
                if low:
                    eng.adjoint_precision(low)
                    if anchor is None:
                        anchor = tuple(x.clone() for x in (Lt1, Lt2, G1, G2))
                    else:
                        for dst, src in zip(anchor, (Lt1, Lt2, G1, G2)):
                            dst.copy_(src)
                    anchors += 1
            elif phase == 2:
                kind = 2
                Lt1.sub_(anchor[0])
                Lt2.sub_(anchor[1])
                eng.adjoint_apply(Lt1, Lt2, G1, G2)
                G1.add_(anchor[2])
                G2.add_(anchor[3])
            else:
                kind = 1
                eng.adjoint_apply(Lt1, Lt2, G1, G2)
            d1, d2, n1, n2 = eng.lambda_update(Lam1, Lam2, G1, G2, T1, g,
                                               beta, damp)
            res = math.sqrt(d1) / nl1 + math.sqrt(d2) / nl2
            nl1 = math.sqrt(n1) + 0.1
            nl2 = math.sqrt(n2) + 0.1
            logging.info(' %2d  %.10f' % (i + 1, res))
            if history is not None:
                history.append(res)
            i += 1
            steps[kind] += 1
            if res < best:
                best, flat = res, 0
            else:
                flat += 1
            if kind == 3:
                converged = res < thresh
                # an FP64 step that does not converge anchors the next ones
                # (after four anchors the iteration stays FP64)
                phase = 2 if (low and not converged and anchors < 4) else 3
            elif phase == 1 and (res < switch or (flat >= 3 and res < 1e-3)):
                phase = 2
                logging.info("Lambda: FP64 anchor at iteration %d" % (i + 1))
            elif phase == 2 and res < thresh:
                phase = 3
        if stats is not None:
            stats["steps"] = steps
        if not converged:
            logging.warning("CCSD Lambda-equations did not converge!")
        logging.info("Total CCSD Lambda time: %f s" % (time.time() - tbeg))
    finally:
        if low:
            eng.adjoint_precision("fp64")
        eng.adjoint_end()
    return Lam1, Lam2, converged, i


def _response_fock(cc, eng, eri, fock_dev, en, fo, fv, dx):
    """(dfock, dE1): the change of the bare Fock matrix and of the
    first-order grand potential when x_p = beta (e_p - mu) moves by dx
    (kelvin/ccsd.py:1434-1492 through g_fock_d_den and g_mp1_den).  The
    uniform electron gas (whose Fock matrix is its mean field) takes it on
    the device from the same integrals; other systems from their
    g_fock_d_den / g_mp1_den."""
    w = -fo * fv * dx
    if hasattr(cc.sys, "g_device_integrals"):
        dfock, _ = eng.mean_field(numpy.zeros_like(en), eri, w)
        jk = torch.diagonal(fock_dev).cpu().numpy() - en
        return dfock, float(numpy.dot(w, jk))
    Fd = numpy.asarray(cc.sys.g_fock_d_den(), dtype=numpy.float64)
    dfock = -numpy.einsum('pqk,k->pq', Fd, dx)
    ron1 = numpy.asarray(cc.sys.g_mp1_den(), dtype=numpy.float64)
    return dfock, float(numpy.dot(ron1 / cc.beta, dx))


def ft_1rdm(cc):
    """The unrelaxed one-particle density of a converged ``ccsd``
    (kelvin/ccsd.py:1333-1387 _g_ft_1rdm): solves the lambda equations if
    they have not been (kelvin calls _ft_ccsd_lambda there), forms the
    response densities on the device (``FTEngine.density1``: kelvin's
    ft_cc_equations.ccsd_1rdm) and sets dia, dba, dji, dai, their
    normal-ordered ndia, ndba, ndji, ndai (scaled by the square roots of the
    occupation numbers) and n1rdm = (ndia + ndba + ndji + ndai) / beta, as
    numpy arrays.  For 'u' orbitals (solved in their spin-orbital form,
    alpha orbitals first) each of these is the pair of its alpha and beta
    blocks, as kelvin/ccsd.py:1579-1665 _u_ft_1rdm sets them (the blocks
    that mix the spins vanish: the density conserves spin)."""
    if cc.beta_max != cc.beta:
        raise Exception("the 1-RDM needs beta_max == beta")
    if cc.athresh > 0.0:
        raise NotImplementedError("the 1-RDM of an active space (athresh > 0) "
                                  "is not built")
    if cc.L2 is None:
        cc.compute_ESN()
    eng = cc.engine
    if eng is None:
        raise Exception("the 1-RDM needs the converged solve's engine")
    beta, mu, g = cc.beta, cc.mu, cc.g
    eng.set_grid(cc.ti, cc.G)
    Lam1, Lam2 = eng.lambda_from_kelvin(cc.L1, cc.L2, g, beta)
    Lt1, Lt2 = eng.empty_amplitudes()
    eng.integrate_adjoint(Lam1, Lam2, Lt1, Lt2)
    del Lam1, Lam2
    pia, pba, pji, pai = (x.cpu().numpy() for x in
                          eng.density1(cc.T1, cc.T2, Lt1, Lt2, g, beta))
    del Lt1, Lt2
    en = numpy.asarray(cc.sys.g_energies_tot(), dtype=numpy.float64)
    sfo = numpy.sqrt(ft_utils.ff(beta, en, mu))
    sfv = numpy.sqrt(ft_utils.ffv(beta, en, mu))
    d = (pia, pba, pji, pai)
    nd = (pia * sfo[:, None] * sfv[None, :], pba * sfv[:, None] * sfv[None, :],
          pji * sfo[:, None] * sfo[None, :], pai * sfv[:, None] * sfo[None, :])
    n1 = sum(nd) / beta
    if cc.sys.orbtype == 'u':
        na = len(cc.sys.u_energies_tot()[0])

        def spins(x):
            return (x[:na, :na].copy(), x[na:, na:].copy())
        d = tuple(spins(x) for x in d)
        nd = tuple(spins(x) for x in nd)
        n1 = list(spins(n1))
    cc.dia, cc.dba, cc.dji, cc.dai = d
    cc.ndia, cc.ndba, cc.ndji, cc.ndai = nd
    cc.n1rdm = n1
    return cc.n1rdm


def compute_esn(cc, L1=None, L2=None, gderiv=True):
    """E, S and N of a converged ``ccsd`` (generalised path of
    kelvin/ccsd.py:128-214); sets the attributes the reference sets (E, S,
    N and their 0 / 1 / cc parts, L1, L2 in kelvin's layout) and returns
    (E, S, N).  gderiv=False replaces the derivative of the quadrature by
    the reference's approximation E_cc(tau_last) / beta
    (kelvin/ccsd.py:1262-1288)."""
    eng = cc.engine
    if eng is None:
        raise Exception("compute_ESN needs the converged solve's engine")
    beta, mu = cc.beta, cc.mu
    if cc.beta_max != beta:
        raise Exception("derivatives need beta_max == beta")
    ng = cc.ngrid
    ti, g, G = cc.ti, cc.g, cc.G
    T1, T2 = cc.T1, cc.T2
    en = numpy.asarray(cc.sys.g_energies_tot(), dtype=numpy.float64)
    n = en.shape[0]
    fo = ft_utils.ff(beta, en, mu)
    fv = ft_utils.ffv(beta, en, mu)
    iocc = ivir = None
    if cc.athresh > 0.0:
        iocc = [i for i, x in enumerate(fo) if x > cc.athresh]
        ivir = [i for i, x in enumerate(fv) if x > cc.athresh]
    t0 = time.time()
    eng.set_precision("fp64")
    eng.set_grid(ti, G)
    skip = zero_slices(eng, T1, T2, G)
    # the residual at the solution (for the grid derivative)
    Tb1 = Tb2 = None
    if gderiv:
        Tb1, Tb2 = eng.empty_amplitudes()
        eng.residual(T1, T2, Tb1, Tb2, skip)
    conv = {"tconv": cc.tconv, "max_iter": cc.max_iter, "damp": cc.damp}
    cc.lambda_history = []
    # a refined bf16x3 forward solve takes the refined lambda iteration
    low = ("bf16x3" if cc.precision == "bf16x3" and cc.refine is True
           else None)
    lstats = {}
    Lam1, Lam2, lconv, lit = solve_lambda(eng, T1, T2, g, beta, conv, L1=L1,
                                          L2=L2, history=cc.lambda_history,
                                          low=low, stats=lstats)
    cc.lambda_converged, cc.lambda_iter = lconv, lit
    cc.lambda_phases = lstats.get("steps")
    cc.adjoint_slices_kept = eng.adjoint_slices_kept
    t_lambda = time.time() - t0
    Omega_cc = cc.Gcc
    # the quadrature's dependence on beta
    if gderiv:
        gd, Gd = quadrature.d_ft_quad(ng, beta, cc.quad)
        dgrid = eng.energy(T1, T2, gd, beta, qterm=True)
        dgrid += eng.grid_derivative(Lam1, Lam2, Tb1, Tb2, Gd, beta)
        del Tb1, Tb2
    else:
        e_last = numpy.zeros(ng)
        e_last[-1] = 1.0
        dgrid = eng.energy(T1, T2, e_last, beta, qterm=True)
    # the integrals' dependence on (beta, mu), contracted with Int^T Lam
    Lt1, Lt2 = eng.empty_amplitudes()
    eng.integrate_adjoint(Lam1, Lam2, Lt1, Lt2)
    cc.L1, cc.L2 = eng.lambda_to_kelvin(Lam1, Lam2, g, beta)
    del Lam1, Lam2
    if hasattr(cc.sys, "g_device_integrals"):
        fock, eri, _ = cc.sys.g_device_integrals(eng)
    else:
        fock = eng.to_device(cc.sys.g_fock_tot())
        eri = eng.to_device(cc.sys.g_aint_tot())
    R1, R2 = eng.empty_amplitudes()

    def h_part(dx):
        dfock, dE1 = _response_fock(cc, eng, eri, fock, en, fo, fv, dx)
        eng.set_system_response(en, fock, dfock, eri, dx, beta, mu, iocc,
                                ivir)
        eng.residual(T1, T2, R1, R2, skip)
        dcc = (eng.energy(T1, T2, g, beta, qterm=True) + eng.dot(Lt1, R1)
               + eng.dot(Lt2, R2))
        return dcc, dE1

    try:
        dcc_mu, dE1_mu = h_part(-beta * numpy.ones(n))
        dcc_beta, dE1_beta = h_part(en - mu)
    finally:
        eng.set_system(en, fock, eri, beta, mu, iocc=iocc, ivir=ivir)
        eng.set_grid(ti, G)
    del R1, R2, Lt1, Lt2, fock, eri
    Ncc = -dcc_mu
    Bcc = dcc_beta - Omega_cc / beta + dgrid
    Ecc = beta * Bcc + mu * Ncc + Omega_cc
    N1 = -dE1_mu
    E1 = beta * dE1_beta + mu * N1 + cc.G1
    N0 = fo.sum()
    E0 = beta * ft_utils.dGP0(beta, en, mu).sum() + mu * N0 + cc.G0
    cc.N0, cc.N1, cc.Ncc = N0, N1, Ncc
    cc.N = N0 + N1 + Ncc
    cc.E0, cc.E1, cc.Ecc = E0, E1, Ecc
    cc.E = E0 + E1 + Ecc
    cc.S = -beta * (cc.Gtot - cc.E + mu * cc.N)
    cc.S0 = -beta * (cc.G0 - cc.E0 + mu * cc.N0)
    cc.S1 = -beta * (cc.G1 - cc.E1 + mu * cc.N1)
    cc.Scc = cc.S - cc.S0 - cc.S1
    cc.time_lambda = t_lambda
    cc.time_esn = time.time() - t0
    logging.info("Total derivative time: %f s" % cc.time_esn)
    return cc.E, cc.S, cc.N


class GrandPotential(torch.autograd.Function):
    """Omega(beta, mu) of FT-CCSD with a hand-written backward: the forward
    runs ``ccsd(make_system(T, mu), T=1/beta, mu=mu, **kw).run()``, the
    backward the adjoint of its converged fixed point (``compute_esn``):
    dOmega/dmu = -N and dOmega/dbeta = (E - mu N - Omega) / beta.  The
    solve's amplitudes are the only state kept between the two (the
    adjoint's intermediates are rebuilt from them, grid point by grid
    point, as its cache allows).  beta and mu are 0-dimensional float64
    tensors."""

    @staticmethod
    def forward(ctx, beta, mu, make_system, kw):
        from .ccsd import ccsd
        b, m = float(beta), float(mu)
        T = 1.0 / b
        cc = ccsd(make_system(T, m), T=T, mu=m, **kw)
        om, _ = cc.run()
        ctx.cc = cc
        return torch.tensor(om, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        cc = ctx.cc
        E, S, N = compute_esn(cc)
        beta, mu = cc.beta, cc.mu
        d_beta = (E - mu * N - cc.Gtot) / beta
        d_mu = -N
        cc.engine = None
        ctx.cc = None
        return (grad * d_beta, grad * d_mu, None, None)


def grand_potential(make_system, beta, mu, **kw):
    """Omega as a differentiable function of beta and mu (tensors)."""
    return GrandPotential.apply(beta, mu, make_system, kw)
