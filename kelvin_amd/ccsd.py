"""Finite-temperature CCSD driver with the reference's user API.

``ccsd(sys, T=, mu=, ngrid=, ...)`` / ``run()`` mirror kelvin/ccsd.py:43-126
and the generalised-orbital solver ``_ft_ccsd`` (kelvin/ccsd.py:595-712):
grid, Fermi-weighted integrals, MP2-level guess, damped fixed point,
Omega = sum GP0 + const + E1 + Omega_cc.  All amplitude work runs on the GPU
(libftcc); ``self.T1`` / ``self.T2`` are torch device tensors.

Extensions over the reference (all optional):
  precision   'fp64' (default), 'fp32' or 'bf16x3' for the O(n^6) and
              O(n^5) contractions; Fermi factors, propagators, quadrature,
              damping, norms and energies stay FP64.
  refine      with a mixed precision, continue in FP64 after convergence
              until the FP64 iteration converges (default True).
"""
import logging
import time

import numpy
import torch

from . import cc_utils
from . import ft_utils
from . import quadrature
from .engine import FTEngine
from .ov_blocks import EngineDenominators, engine_eri, engine_fock


class ccsd(object):
    """Coupled cluster singles and doubles (CCSD) driver at finite T."""

    def __init__(self, sys, T=0.0, mu=0.0, iprint=0, singles=True, econv=1e-8,
                 tconv=None, max_iter=40, damp=0.0, ngrid=10, realtime=False,
                 athresh=0.0, quad='lin', rt_iter="all", precision="fp64",
                 refine=True, refine_max_iter=None):
        self.T = T
        self.mu = mu
        self.finite_T = False if T == 0 else True
        self.iprint = iprint
        self.singles = singles
        self.econv = econv
        self.tconv = tconv if tconv is not None else 1000.0 * econv
        self.max_iter = max_iter
        self.damp = damp
        self.ngrid = ngrid
        self.realtime = realtime
        self.athresh = athresh
        self.quad = quad
        self.rt_iter = rt_iter
        self.precision = precision
        self.refine = refine
        self.refine_max_iter = refine_max_iter
        if self.finite_T:
            self.realtime = True
        if not sys.verify(self.T, self.mu):
            raise Exception("Sytem temperature inconsistent with CC temp")
        self.beta = None
        self.ti = None
        self.g = None
        self.G = None
        if self.realtime:
            if self.finite_T:
                self.beta = 1.0 / T
                self.beta_max = self.beta
            else:
                self.beta_max = 80.0
            self.ti, self.g, self.G = quadrature.ft_quad(self.ngrid,
                                                         self.beta_max,
                                                         self.quad)
        self.sys = sys
        self.T1 = None
        self.T2 = None
        self.L1 = None
        self.L2 = None
        self.engine = None
        self.history = []

    def run(self, T1=None, T2=None):
        """Run CCSD; returns (Omega_total, Omega_cc)."""
        if not self.finite_T:
            raise NotImplementedError(
                "zero-temperature CCSD is not part of the GPU path")
        logging.info('Running CCSD at an electronic temperature of %f K'
                     % ft_utils.HtoK(self.T))
        if not self.singles:
            raise NotImplementedError("CCD is not built for the GPU path")
        if self.rt_iter[0] != 'a' and T2 is None:
            raise NotImplementedError(
                "rt_iter='point' is not built for the GPU path")
        return self._ft_ccsd(T1in=T1, T2in=T2)

    def compute_ESN(self, L1=None, L2=None, gderiv=True):
        raise NotImplementedError(
            "E/S/N (lambda equations) are not built for the GPU path yet")

    def _ft_ccsd(self, T1in=None, T2in=None):
        ng = self.ngrid
        ti, g, G = self.ti, self.g, self.G
        beta, mu = self.beta, self.mu
        en = numpy.asarray(self.sys.g_energies_tot(), dtype=numpy.float64)
        n = en.shape[0]
        En = self.sys.const_energy()
        if self.athresh > 0.0:
            fo = ft_utils.ff(beta, en, mu)
            fv = ft_utils.ffv(beta, en, mu)
            iocc = [i for i, x in enumerate(fo) if x > self.athresh]
            ivir = [i for i, x in enumerate(fv) if x > self.athresh]
            logging.info("FT-CCSD orbital info:")
            logging.info('  nocc: {:d}'.format(len(iocc)))
            logging.info('  nvir: {:d}'.format(len(ivir)))
            logging.info('  nact: {:d}'.format(len(iocc) + len(ivir) - n))
        else:
            iocc = ivir = None
        no = n if iocc is None else len(iocc)
        nv = n if ivir is None else len(ivir)
        prec0 = self.precision
        eng = FTEngine(no, nv, ng, precision=prec0)
        self.engine = eng
        t0 = time.time()
        G0 = eng.set_system(en, self.sys.g_fock_tot(), self.sys.g_aint_tot(),
                            beta, mu, iocc=iocc, ivir=ivir)
        E0 = G0 + En
        E1 = self.sys.get_mp1()
        E01 = E0 + E1
        eng.set_grid(ti, G)
        logging.info("integrals on device: %.3f s (%.2f GB held)"
                     % (time.time() - t0, eng.workspace_bytes() / 1e9))
        F, I = engine_fock(eng), engine_eri(eng)
        D1 = EngineDenominators(eng, 1)
        D2 = EngineDenominators(eng, 2)
        if T1in is not None and T2in is not None:
            T1old = eng.to_device(T1in).clone()
            T2old = eng.to_device(T2in).clone()
        else:
            T1old, T2old = eng.empty_amplitudes()
            eng.driving_terms(T1old, T2old)
            eng.integrate(T1old, T2old)
        E2 = eng.energy(T1old, T2old, g, self.beta_max, qterm=False)
        logging.info('MP2 Energy: {:.10f}'.format(E2))
        self.mp2 = E2
        conv = {"econv": self.econv, "tconv": self.tconv,
                "max_iter": self.max_iter, "damp": self.damp}
        self.history = []
        Eccn, T1, T2 = cc_utils.ft_cc_iter(
            "CCSD", T1old, T2old, F, I, D1, D2, g, G, self.beta_max, ng, ti,
            self.iprint, conv, history=self.history)
        if prec0 != "fp64" and self.refine:
            logging.info("FP64 refinement after %s iterations" % prec0)
            eng.set_precision("fp64")
            conv_r = dict(conv)
            if self.refine_max_iter is not None:
                conv_r["max_iter"] = self.refine_max_iter
            Eccn, T1, T2 = cc_utils.ft_cc_iter(
                "CCSD", T1, T2, F, I, D1, D2, g, G, self.beta_max, ng, ti,
                self.iprint, conv_r, history=self.history)
        self.T1 = T1
        self.T2 = T2
        self.G0 = E0
        self.G1 = E1
        self.Gcc = Eccn
        self.Gtot = E0 + E1 + Eccn
        return (Eccn + E01, Eccn)
