"""numpy FP64 restatement of kelvin's imaginary-time propagation (TD-CCSD).

TEST INFRASTRUCTURE ONLY: this is the oracle (see oracle/ftcc.py).  The
product package ``kelvin_amd`` never imports it.

Restates kelvin/td_ccsd.py:583-681 (``TDCCSD._ccsd``, generalised orbitals,
finite temperature), the explicit propagators of kelvin/propagation.py:5-52
(rk1, rk2, rk4) and its implicit ones (be, cn, am2, :55-128).  The right-hand side
-D T - F.vo / -D T - I.vvoo minus the CCSD residual with fac=-1
(td_ccsd.py:646-652) is, in this oracle's conventions, Tbar(T) - D T with
Tbar = -stanton_rhs.

Pinned (tests/test_oracle.py):
  * rk1/rk2/rk4 on the linear, interaction-free problem whose exact solution
    is T1(tau) = -F.vo (1 - exp(-D1 tau)) / D1: the error falls by 2, 4 and
    16 when the step halves (a wrong tableau coefficient breaks the order);
    be, cn and am2 (inner iteration converged tightly) by 2, 4 and 4;
  * the propagated grand potential against the fixed-point FT-CCSD one on
    fine grids (the reference's own cross-check, kelvin/tests/
    test_td_ccsd.py:12-75).
"""
import numpy

from .ftcc import stanton_rhs


def td_rhs(F, I, D1, D2, t1, t2, singles=True):
    """kelvin/td_ccsd.py:646-652 for one state (t1 [nv,no], t2 [nv,nv,no,no])."""
    r1, r2 = stanton_rhs(F, I, t1, t2)
    k1 = -D1 * t1 - r1
    k2 = -D2 * t2 - r2
    if not singles:
        k1 = numpy.zeros_like(k1)
    return k1, k2


def rk_step(order, h, var, rhs):
    """kelvin/propagation.py:5-52: the increment (not the new state)."""
    if order == 1:
        return [h * x for x in rhs(var)]
    if order == 2:
        k1 = [h * x for x in rhs(var)]
        k2 = rhs([x + y for x, y in zip(k1, var)])
        return [0.5 * (x + h * y) for x, y in zip(k1, k2)]
    if order == 4:
        k1 = [h * x for x in rhs(var)]
        k2 = [h * x for x in rhs([x + 0.5 * y for x, y in zip(var, k1)])]
        k3 = [h * x for x in rhs([x + 0.5 * y for x, y in zip(var, k2)])]
        k4 = [h * x for x in rhs([x + y for x, y in zip(var, k3)])]
        return [(a + 2.0 * b + 2.0 * c + d) / 6.0
                for a, b, c, d in zip(k1, k2, k3, k4)]
    raise ValueError("order must be 1, 2 or 4")


def implicit_step(scheme, h, var, rhs, max_iter, alpha, thresh):
    """kelvin/propagation.py:55-128: be, cn (backward-Euler iteration from
    the explicit predictor, then the average with it) and am2 (trapezoid).
    The damped fixed point d <- (1 - alpha) d_new + alpha d stops when
    sum ||d_new - d|| / (||d_new|| + 0.01) < thresh.  Returns (increment,
    iterations); raises like the reference when it does not converge."""
    k1 = [h * x for x in rhs(var)]
    dold = k1
    for it in range(max_iter):
        k = rhs([x + y for x, y in zip(var, dold)])
        if scheme == "am2":
            d = [0.5 * (h * x + y) for x, y in zip(k, k1)]
        else:
            d = [h * x for x in k]
        error = sum(numpy.linalg.norm(x - y) / (numpy.linalg.norm(x) + 0.01)
                    for x, y in zip(d, dold))
        dold = [(1.0 - alpha) * x + alpha * y for x, y in zip(d, dold)]
        if error < thresh:
            break
    else:
        raise Exception("%s: Failed to compute implicit step"
                        % ("AM2" if scheme == "am2" else "BE"))
    if scheme == "cn":
        return [0.5 * (x + y) for x, y in zip(dold, k1)], it + 1
    return dold, it + 1


def cc_energy(t1, t2, fov, oovv):
    """CC energy of one state (cqcpy cc_energy, used at td_ccsd.py:671)."""
    return (numpy.einsum('ia,ai->', fov, t1)
            + 0.25 * numpy.einsum('ijab,abij->', oovv, t2)
            + 0.5 * numpy.einsum('ijab,ai,bj->', oovv, t1, t1))


def td_ccsd(F, I, D1, D2, ti, g, beta, order, singles=True, implicit=None):
    """kelvin/td_ccsd.py:644-681: propagate from T = 0 over the grid ti and
    accumulate Omega_cc = sum_{i>=1} g_i E(T(tau_i)) / beta.  Returns
    (Omega_cc, t1, t2) with the final amplitudes.  ``order`` 1, 2, 4 selects
    rk1/rk2/rk4; ``implicit`` = (scheme, max_iter, damp, thresh) selects
    be/cn/am2 instead."""
    nv, no = F.vo.shape
    t1 = numpy.zeros((nv, no))
    t2 = numpy.zeros((nv, nv, no, no))

    def rhs(var):
        return td_rhs(F, I, D1, D2, var[0], var[1], singles)

    Ecc = 0.0
    for i in range(1, len(ti)):
        h = ti[i] - ti[i - 1]
        if implicit is None:
            d1, d2 = rk_step(order, h, [t1, t2], rhs)
        else:
            scheme, mi, alpha, thresh = implicit
            (d1, d2), _ = implicit_step(scheme, h, [t1, t2], rhs, mi, alpha,
                                        thresh)
        if singles:
            t1 = t1 + d1
        t2 = t2 + d2
        Ecc += g[i] * cc_energy(t1, t2, F.ov, I.oovv) / beta
    return Ecc, t1, t2
