"""Imaginary-time propagated FT-CCSD (TD-CCSD) with the reference's user API.

``TDCCSD(sys, prop, T=, mu=, ngrid=, ...).run()`` mirrors kelvin/td_ccsd.py:
74-158 (construction, run) and the generalised-orbital propagation
``_ccsd`` (kelvin/td_ccsd.py:583-681): starting from T = 0 at tau = 0, the
amplitudes are propagated over the quadrature grid with
dT/dtau = -D T - (F.vo, I.vvoo) - R_CCSD(T) and the grand potential
accumulates Omega_cc = sum_{i>=1} g_i E_CC(T(tau_i)) / beta.  Every step runs
on the GPU (``ftcc_td_step``: the residual engine plus one fused kernel per
Runge-Kutta stage; ``ftcc_td_implicit_step``: the damped inner iteration of
the implicit schemes, one fused update-and-norm kernel per pass); the
propagators are kelvin/propagation.py's explicit rk1, rk2, rk4 (:5-52) and
implicit be, cn, am2 (:55-128, with the ``max_iter``, ``damp`` and
``thresh`` entries of ``prop``).  ``compute_ESN()`` gives E, S and N
(kelvin/td_ccsd.py:160-235) as derivatives of the propagated grand
potential.

It is an independent check of the fixed-point solver: both converge to the
same Omega_cc as the grid is refined (kelvin/tests/test_td_ccsd.py).
"""
import logging

import numpy

from . import ft_utils
from . import quadrature
from .engine import FTEngine

_ORDERS = {"rk1": 1, "rk2": 2, "rk4": 4}
_IMPLICIT = ("be", "cn", "am2")


class TDCCSD(object):
    """Time-dependent (imaginary-time propagated) CCSD solver."""

    def __init__(self, sys, prop, T=0.0, mu=0.0, iprint=0, singles=True,
                 ngrid=10, athresh=0.0, fthresh=0.0, quad='lin',
                 saveT=False, saveL=False, tmem="mem", scratch="",
                 precision="fp64"):
        self.T = T
        self.mu = mu
        self.finite_T = False if T == 0 else True
        self.iprint = iprint
        self.singles = singles
        self.ngrid = ngrid
        self.athresh = athresh
        self.fthresh = fthresh
        self.quad = quad
        self.prop = prop
        self.saveT = saveT
        self.saveL = saveL
        self.tmem = tmem
        self.scratch = scratch
        self.precision = precision
        if not sys.verify(self.T, self.mu):
            raise Exception("Sytem temperature inconsistent with CC temp")
        self.beta = 1.0 / T if self.finite_T else 80
        self.ti, self.g, _ = quadrature.ft_quad(self.ngrid, self.beta,
                                                self.quad)
        self.sys = sys
        self.T1 = [None] * ngrid
        self.T2 = [None] * ngrid
        self.G0 = self.G1 = self.Gcc = self.Gtot = None
        self.E = self.S = self.N = None
        self.engine = None

    def run(self, response=None):
        """Run the propagation; returns (Omega_total, Omega_cc)."""
        if not self.finite_T:
            raise NotImplementedError(
                "zero-temperature TD-CCSD is not part of the GPU path")
        tprop = self.prop.get("tprop")
        if tprop not in _ORDERS and tprop not in _IMPLICIT:
            raise Exception("Unrecognized propagation scheme: " + str(self.prop))
        logging.info('Running CCSD at an electronic temperature of %f K'
                     % ft_utils.HtoK(self.T))
        return self._ccsd(tprop)

    def compute_ESN(self, delta=1e-4):
        """E, S and N with kelvin's attributes (E0/E1/Ecc, S0/S1/Scc,
        N0/N1/Ncc and their sums; kelvin/td_ccsd.py:160-235): the zero
        order parts in closed form (kelvin/td_ccsd.py:200-205), the first
        order and CC parts as N = -dOmega/dmu and S = -dOmega/dT by central
        differences of propagations at (T, mu +- delta) and (T +- delta, mu)
        on grids scaled with beta, and E = Omega + T S + mu N for every
        part.  A propagation has no convergence test, so the differences
        carry only the O(delta^2) truncation: the result is the derivative
        of the propagated grand potential itself.  Kelvin obtains the same
        derivatives by propagating the lambda amplitudes backwards and
        contracting response densities, with an approximate grid
        derivative; the two agree with the fixed-point solver's E, S and N
        to the tolerance of kelvin/tests/test_td_ccsd_ESN.py:42-75.  The
        system must derive its (T, mu)-dependent quantities from its ``T``
        and ``mu`` attributes (true of the systems in this package)."""
        if not self.finite_T:
            raise NotImplementedError(
                "zero-temperature TD-CCSD is not part of the GPU path")
        from .ccsd import system_at
        if self.Gtot is None:
            self.run()
        T, mu, beta = self.T, self.mu, self.beta
        en = numpy.asarray(self.sys.g_energies_tot(), dtype=numpy.float64)
        fo = ft_utils.ff(beta, en, mu)
        N0 = fo.sum()
        E0 = beta * ft_utils.dGP0(beta, en, mu).sum() + mu * N0 + self.G0

        def parts(Ts, mus):
            """(Omega_1, Omega_cc) of the propagation at (Ts, mus)."""
            td = TDCCSD(system_at(self.sys, Ts, mus), self.prop, T=Ts,
                        mu=mus, iprint=self.iprint, singles=self.singles,
                        ngrid=self.ngrid, athresh=self.athresh,
                        fthresh=self.fthresh, quad=self.quad,
                        precision=self.precision)
            td.run()
            return td.G1, td.Gcc

        # each neighbouring propagation holds an engine of its own
        self.engine = None
        e1p, ccp = parts(T, mu + delta)
        e1m, ccm = parts(T, mu - delta)
        e1f, ccf = parts(T + delta, mu)
        e1b, ccb = parts(T - delta, mu)
        N1 = -(e1p - e1m) / (2 * delta)
        Ncc = -(ccp - ccm) / (2 * delta)
        S1 = -(e1f - e1b) / (2 * delta)
        Scc = -(ccf - ccb) / (2 * delta)
        self.N0, self.N1, self.Ncc = N0, N1, Ncc
        self.N = N0 + N1 + Ncc
        self.E0 = E0
        self.E1 = self.G1 + T * S1 + mu * N1
        self.Ecc = self.Gcc + T * Scc + mu * Ncc
        self.E = self.E0 + self.E1 + self.Ecc
        self.S = -beta * (self.Gtot - self.E + mu * self.N)
        self.S0 = -beta * (self.G0 - self.E0 + mu * self.N0)
        self.S1 = -beta * (self.G1 - self.E1 + mu * self.N1)
        self.Scc = self.S - self.S0 - self.S1
        return self.E, self.S, self.N

    def full_1rdm(self, relax=False):
        """kelvin/td_ccsd.py:1996-2031: the densities of the backward lambda
        propagation, which is not built (ccsd.full_1rdm gives the
        fixed-point density of the same system)."""
        raise NotImplementedError(
            "TD-CCSD densities (backward lambda propagation) are not built")

    def full_2rdm(self, relax=False):
        """kelvin/td_ccsd.py:2032-2062: not built (see full_1rdm)."""
        if relax:
            raise Exception("Rexalex 2-RDM is not implemented")
        raise NotImplementedError(
            "TD-CCSD densities (backward lambda propagation) are not built")

    def _active(self, en):
        """kelvin/td_ccsd.py:17-72 ('g' branch): orbitals with occupation
        above athresh and vacancy above fthresh, and vice versa."""
        fo = ft_utils.ff(self.beta, en, self.mu)
        fv = ft_utils.ffv(self.beta, en, self.mu)
        iocc = [i for i, x in enumerate(fo)
                if x > self.athresh and fv[i] > self.fthresh]
        ivir = [i for i, x in enumerate(fv)
                if x > self.athresh and fo[i] > self.fthresh]
        logging.info("FT-CCSD orbital info:")
        logging.info("  nocc: {:d}".format(len(iocc)))
        logging.info("  nvir: {:d}".format(len(ivir)))
        return iocc, ivir

    def _ccsd(self, tprop):
        beta, mu = self.beta, self.mu
        ng, ti, g = self.ngrid, self.ti, self.g
        en = numpy.asarray(self.sys.g_energies_tot(), dtype=numpy.float64)
        n = en.shape[0]
        iocc, ivir = self._active(en)
        full = len(iocc) == n and len(ivir) == n
        eng = FTEngine(len(iocc), len(ivir), 1, precision=self.precision)
        self.engine = eng
        if hasattr(self.sys, "g_device_integrals"):
            fock, eri, E1 = self.sys.g_device_integrals(eng)
        else:
            fock, eri = self.sys.g_fock_tot(), self.sys.g_aint_tot()
            E1 = self.sys.get_mp1()
        G0 = eng.set_system(en, fock, eri, beta, mu,
                            iocc=None if full else iocc,
                            ivir=None if full else ivir)
        del fock, eri
        E0 = G0 + self.sys.const_energy()
        E01 = E0 + E1
        # one state propagated; the grid of the context only sizes it
        eng.set_grid(numpy.zeros(1), numpy.zeros((1, 1)))
        T1, T2 = eng.empty_amplitudes()
        T1.zero_()
        T2.zero_()
        Eccn = 0.0
        if self.saveT:
            self.T1[0], self.T2[0] = T1[0].clone(), T2[0].clone()
        for i in range(1, ng):
            h = ti[i] - ti[i - 1]
            if tprop in _ORDERS:
                eng.td_step(T1, T2, h, order=_ORDERS[tprop],
                            singles=self.singles)
            else:
                eng.td_implicit_step(T1, T2, h, tprop, self.prop["max_iter"],
                                     self.prop["damp"], self.prop["thresh"],
                                     singles=self.singles)
            Eccn += eng.energy(T1, T2, numpy.array([g[i]]), beta, qterm=True)
            if self.saveT:
                self.T1[i], self.T2[i] = T1[0].clone(), T2[0].clone()
        self.G0 = E0
        self.G1 = E1
        self.Gcc = Eccn
        self.Gtot = E0 + E1 + Eccn
        if not self.saveT:
            self.T1[ng - 1], self.T2[ng - 1] = T1[0], T2[0]
        return (Eccn + E01, Eccn)
