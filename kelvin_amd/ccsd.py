"""Finite-temperature CCSD solver with the reference's user API.

``ccsd(sys, T=, mu=, ngrid=, ...)`` / ``run()`` mirror kelvin/ccsd.py:43-126
and the generalised-orbital solver ``_ft_ccsd`` (kelvin/ccsd.py:595-712):
grid, Fermi-weighted integrals, MP2-level guess, damped fixed point,
Omega = sum GP0 + const + E1 + Omega_cc.  All amplitude work runs on the GPU
(libftcc); ``self.T1`` / ``self.T2`` are torch device tensors.

Extensions over the reference (all optional):
  precision   'fp64' (default), 'fp32' or 'bf16x3' for the O(n^6) and
              O(n^5) contractions; Fermi factors, propagators, quadrature,
              damping, norms and energies stay FP64.
  refine      with a mixed precision: True (default) iterates in the low
              precision until the change reaches its rounding floor, then
              with the FP64 defect (FP64 minus low-precision residual at
              that iterate) added to every residual, and finishes with FP64
              steps, so that convergence is certified by an FP64 step
              (``cc_utils.ft_cc_iter_refined``); "fp64" continues with
              plain FP64 iterations after the low-precision solve
              converges; False stops at the low-precision fixed point.
  graph       run the solve on a stream of its own with every fixed-point
              step replayed as a HIP graph (launch overhead of small
              systems) (default False).  Without a guess (T1/T2 on
              entry) the solver's own iterates are trusted to conserve
              spin and the per-step check of the amplitudes is skipped.
  chunk       grid points per residual chunk (default: chosen from the
              workspace budget; ``kelvin_amd.batch.autotune`` times the
              choices).
  gemm_plan   (tile_policy, split_target) of the generic GEMMs (default
              the library's; ``kelvin_amd.batch.autotune`` times them).
  profile     time every kernel class of the solve with HIP events;
              ``kernel_classes`` then holds {class: (launches, ms, flops,
              bytes)} of the whole solve (default False).
"""
import copy
import logging
import time

import numpy
import torch

from . import cc_utils
from . import ft_utils
from . import quadrature
from .engine import FTEngine
from .ov_blocks import EngineDenominators, engine_eri, engine_fock


def system_at(sys, T, mu):
    """The system at another (T, mu): a shallow copy whose temperature and
    chemical potential are replaced (the systems of this package derive
    their (T, mu)-dependent quantities from these attributes)."""
    s = copy.copy(sys)
    s.T = T
    s.mu = mu
    if hasattr(s, "beta"):
        s.beta = 1.0 / T
    if not s.verify(T, mu):
        raise Exception("system cannot be moved to T=%r, mu=%r" % (T, mu))
    return s


class ccsd(object):
    """Coupled cluster singles and doubles (CCSD) solver at finite T."""

    def __init__(self, sys, T=0.0, mu=0.0, iprint=0, singles=True, econv=1e-8,
                 tconv=None, max_iter=40, damp=0.0, ngrid=10, realtime=False,
                 athresh=0.0, quad='lin', rt_iter="all", precision="fp64",
                 refine=True, refine_max_iter=None, graph=False,
                 chunk=None, gemm_plan=None, profile=False):
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
        self.graph = graph
        self.chunk = chunk
        self.gemm_plan = gemm_plan
        self.profile = profile
        self.kernel_classes = None
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
        # pieces of the unrelaxed 1-RDM and the 1-RDM itself
        # (kelvin/ccsd.py:84-103)
        self.dia = self.dba = self.dji = self.dai = None
        self.ndia = self.ndba = self.ndji = self.ndai = None
        self.n1rdm = None
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
        if not self.graph:
            return self._ft_ccsd(T1in=T1, T2in=T2)
        # graph steps need a stream of their own (not the null stream)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            out = self._ft_ccsd(T1in=T1, T2in=T2)
        torch.cuda.current_stream().wait_stream(stream)
        return out

    def compute_ESN(self, L1=None, L2=None, gderiv=True):
        """Energy, entropy and particle number (kelvin/ccsd.py:128-214).

        The derivatives of the discretised grand potential with beta and mu
        come from the adjoint of the converged fixed point
        (``kelvin_amd.adjoint.compute_esn``): lambda amplitudes from
        kelvin's damped iteration (L1 / L2 in kelvin's layout start it, as
        in kelvin/ccsd.py:926-946), the response of the Fermi-weighted
        integrals, of the Fock matrix and of the grid, all on the GPU.
        Sets E, S, N, their 0 / 1 / cc parts and L1, L2 as the reference
        does; gderiv=False takes the reference's approximate grid
        derivative (kelvin/ccsd.py:1262-1288).  ``compute_ESN_fd`` gives the
        same quantities by finite differences of full solves."""
        if not self.finite_T:
            logging.warning(
                "Computing thermodynamic quantities at zero temperature")
            raise NotImplementedError(
                "zero-temperature CCSD is not part of the GPU path")
        if self.T1 is None or self.engine is None:
            self.run()
        from .adjoint import compute_esn
        return compute_esn(self, L1=L1, L2=L2, gderiv=gderiv)

    def full_1rdm(self, relax=False):
        """The one-particle density matrix (kelvin/ccsd.py:1961-2006): the
        unrelaxed response density n1rdm (``kelvin_amd.adjoint.ft_1rdm``,
        formed on the first call, solving the lambda equations first if
        needed) plus the Fermi occupations on the diagonal; for 'u' systems
        the [alpha, beta] pair.  The relaxed form (relax=True: the
        occupation and orbital-energy relaxation of kelvin's _grel_ft_1rdm
        and _urel_ft_1rdm) is not built."""
        orbtype = self.sys.orbtype
        if orbtype not in ('g', 'u'):
            raise Exception("orbital type " + orbtype +
                            " is not implemented for 1rdm")
        if relax:
            raise NotImplementedError("the relaxed 1-RDM is not built")
        if self.n1rdm is None:
            from .adjoint import ft_1rdm
            ft_1rdm(self)
        if orbtype == 'u':
            ea, eb = self.sys.u_energies_tot()
            return [self.n1rdm[0] + numpy.diag(ft_utils.ff(self.beta, ea,
                                                           self.mu)),
                    self.n1rdm[1] + numpy.diag(ft_utils.ff(self.beta, eb,
                                                           self.mu))]
        en = self.sys.g_energies_tot()
        return self.n1rdm + numpy.diag(ft_utils.ff(self.beta, en, self.mu))

    def full_2rdm(self, relax=False):
        """kelvin/ccsd.py:2008-2053, the two-particle density: not built (E,
        S, N and the one-particle density do not need it); the reference's
        error for relax=True, NotImplementedError otherwise."""
        if relax:
            raise Exception("Rexalex 2-RDM is not implemented")
        raise NotImplementedError("the two-particle density is not built")

    def compute_ESN_fd(self, delta=1e-4, richardson=False, fd_econv=1e-12):
        """E, S and N by central differences of full GPU solves.

        N = -dOmega/dmu and S = -dOmega/dT from solves at (T, mu +- delta)
        and (T +- delta, mu), each warm-started from the converged
        amplitudes and converged to ``fd_econv``, with the grid rescaled
        with beta as in the reference; the truncation error is O(delta^2)
        (O(delta^4) with ``richardson=True``, which adds four solves at
        2*delta).  The zero order parts are the closed forms of
        kelvin/ccsd.py:175-180; E = Omega + T S + mu N for every part.  The
        system must derive its (T, mu)-dependent quantities from its ``T``
        and ``mu`` attributes (true of the systems in this package).  A
        check of ``compute_ESN``, costing several solves.
        """
        if not self.finite_T:
            raise NotImplementedError(
                "zero-temperature CCSD is not part of the GPU path")
        if self.T1 is None:
            self.run()
        T, mu, beta = self.T, self.mu, self.beta
        en = numpy.asarray(self.sys.g_energies_tot(), dtype=numpy.float64)

        # zero order (closed forms)
        fo = ft_utils.ff(beta, en, mu)
        N0 = fo.sum()
        E0 = beta * ft_utils.dGP0(beta, en, mu).sum() + mu * N0 + self.G0

        def parts(Ts, mus):
            """(E1, Omega_cc) at a neighbouring (T, mu)."""
            sysx = self._system_at(Ts, mus)
            cc = ccsd(sysx, T=Ts, mu=mus, iprint=self.iprint,
                      singles=self.singles, econv=min(self.econv, fd_econv),
                      tconv=self.tconv, max_iter=max(self.max_iter, 200),
                      damp=self.damp, ngrid=self.ngrid, athresh=self.athresh,
                      quad=self.quad, rt_iter=self.rt_iter,
                      precision=self.precision, refine=True)
            warm = self._warm_start(cc)
            cc.run(T1=warm[0], T2=warm[1])
            out = (cc.G1, cc.Gcc)
            del cc
            return out

        def derivs(h):
            e1p, ccp = parts(T, mu + h)
            e1m, ccm = parts(T, mu - h)
            e1f, ccf = parts(T + h, mu)
            e1b, ccb = parts(T - h, mu)
            return (numpy.array([-(e1p - e1m), -(ccp - ccm)]) / (2 * h),
                    numpy.array([-(e1f - e1b), -(ccf - ccb)]) / (2 * h))

        # the neighbouring solves each hold a full engine: release this one
        # (its converged amplitudes stay for the warm starts)
        self.engine = None
        # the base point itself at the finite-difference convergence
        e1c, gcc = parts(T, mu)
        n_h, s_h = derivs(delta)
        if richardson:
            n_2h, s_2h = derivs(2 * delta)
            n_h = (4.0 * n_h - n_2h) / 3.0
            s_h = (4.0 * s_h - s_2h) / 3.0
        N1, Ncc = n_h
        S1, Scc = s_h
        E1 = e1c + T * S1 + mu * N1
        Ecc = gcc + T * Scc + mu * Ncc

        self.N0, self.N1, self.Ncc = N0, N1, Ncc
        self.N = N0 + N1 + Ncc
        self.E0, self.E1, self.Ecc = E0, E1, Ecc
        self.E = E0 + E1 + Ecc
        Gtot = self.G0 + e1c + gcc
        self.S = -beta * (Gtot - self.E + mu * self.N)
        self.S0 = -beta * (self.G0 - self.E0 + mu * self.N0)
        self.S1 = -beta * (e1c - self.E1 + mu * self.N1)
        self.Scc = self.S - self.S0 - self.S1
        return self.E, self.S, self.N

    def _system_at(self, T, mu):
        return system_at(self.sys, T, mu)

    def _warm_start(self, other):
        """Converged amplitudes as the starting point of a neighbouring
        solve, when the active space has the same size."""
        if self.T1 is None:
            return None, None
        en = numpy.asarray(other.sys.g_energies_tot(), dtype=numpy.float64)
        n = en.shape[0]
        if other.athresh > 0.0:
            fo = ft_utils.ff(other.beta, en, other.mu)
            fv = ft_utils.ffv(other.beta, en, other.mu)
            shape = (int((fv > other.athresh).sum()),
                     int((fo > other.athresh).sum()))
        else:
            shape = (n, n)
        if tuple(self.T1.shape[1:]) != shape:
            return None, None
        return self.T1, self.T2

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
        if self.chunk is not None:
            eng.set_chunk(self.chunk)
        if self.gemm_plan is not None:
            eng.set_gemm_plan(*self.gemm_plan)
        if self.graph:
            eng.use_graph = True
        if self.profile:
            eng.profile(True)
            eng.profile_classes()
        # iterates started from the driving terms conserve spin: the
        # per-step check of the amplitudes is needed for a guess only
        eng.set_trusted(T1in is None and T2in is None)
        t0 = time.time()
        if hasattr(self.sys, "g_device_integrals"):
            # integrals built in HBM (the UEG provider)
            fock, eri, E1 = self.sys.g_device_integrals(eng)
        else:
            fock, eri = self.sys.g_fock_tot(), self.sys.g_aint_tot()
            E1 = self.sys.get_mp1()
        G0 = eng.set_system(en, fock, eri, beta, mu, iocc=iocc, ivir=ivir)
        del fock, eri
        E0 = G0 + En
        E01 = E0 + E1
        eng.set_grid(ti, G)
        logging.info("integrals on device: %.3f s (%.2f GB held)"
                     % (time.time() - t0, eng.workspace_bytes() / 1e9))
        F, I = engine_fock(eng), engine_eri(eng)
        D1 = EngineDenominators(eng, 1)
        D2 = EngineDenominators(eng, 2)
        conv = {"econv": self.econv, "tconv": self.tconv,
                "max_iter": self.max_iter, "damp": self.damp}
        if self.rt_iter[0] != 'a' and T2in is None:
            # pointwise schedule (kelvin/ccsd.py:697-700)
            self.history = []
            T1, T2 = cc_utils.ft_cc_iter_extrap(
                "CCSD", F, I, D1, D2, g, G, self.beta_max, ng, ti,
                self.iprint, conv)
            if prec0 != "fp64" and self.refine:
                # pointwise schedule: plain FP64 iterations of the whole grid
                # from amplitudes whose spin-flip images agree only to the
                # low precision's rounding: no trusted promise
                logging.info("FP64 refinement after %s iterations" % prec0)
                eng.set_trusted(False)
                eng.set_precision("fp64")
                conv_r = dict(conv)
                if self.refine_max_iter is not None:
                    conv_r["max_iter"] = self.refine_max_iter
                Eccn, T1, T2 = cc_utils.ft_cc_iter(
                    "CCSD", T1, T2, F, I, D1, D2, g, G, self.beta_max, ng,
                    ti, self.iprint, conv_r, history=self.history)
            Eccn = eng.energy(T1, T2, g, self.beta_max, qterm=True)
            self.T1, self.T2 = T1, T2
            self.G0, self.G1, self.Gcc = E0, E1, Eccn
            self.Gtot = E0 + E1 + Eccn
            return (Eccn + E01, Eccn)
        if self.rt_iter[0] != 'a':
            logging.warning("Converngece scheme ({}) is being ignored."
                            .format(self.rt_iter))
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
        self.history = []
        self.corrections = 0
        self.phase_steps = None
        t_mixed = time.time()
        if prec0 != "fp64" and self.refine is True:
            # low-precision iterations with FP64 defect corrections; the
            # accepted step is an FP64 step
            stats = {}
            Eccn, T1, T2 = cc_utils.ft_cc_iter_refined(
                "CCSD", T1old, T2old, F, I, D1, D2, g, G, self.beta_max, ng,
                ti, self.iprint, conv, prec0, history=self.history,
                stats=stats)
            self.phase_steps = stats["steps"]
            self.corrections = 1 if stats["steps"][2] else 0
        else:
            Eccn, T1, T2 = cc_utils.ft_cc_iter(
                "CCSD", T1old, T2old, F, I, D1, D2, g, G, self.beta_max, ng,
                ti, self.iprint, conv, history=self.history)
        torch.cuda.current_stream(eng.device).synchronize()
        # the solve in the chosen precision, before any continuation in FP64
        self.Gcc_unrefined = Eccn
        self.niter_unrefined = len(self.history)
        self.time_unrefined = time.time() - t_mixed
        self.niter_refine = 0
        self.time_refine = 0.0
        if prec0 != "fp64" and self.refine == "fp64":
            logging.info("FP64 refinement after %s iterations" % prec0)
            t_ref = time.time()
            eng.set_precision("fp64")
            conv_r = dict(conv)
            if self.refine_max_iter is not None:
                conv_r["max_iter"] = self.refine_max_iter
            Eccn, T1, T2 = cc_utils.ft_cc_iter(
                "CCSD", T1, T2, F, I, D1, D2, g, G, self.beta_max, ng, ti,
                self.iprint, conv_r, history=self.history)
            torch.cuda.current_stream(eng.device).synchronize()
            self.niter_refine = len(self.history) - self.niter_unrefined
            self.time_refine = time.time() - t_ref
        if self.profile:
            self.kernel_classes = eng.profile_classes()
            eng.profile(False)
        self.T1 = T1
        self.T2 = T2
        self.G0 = E0
        self.G1 = E1
        self.Gcc = Eccn
        self.Gtot = E0 + E1 + Eccn
        return (Eccn + E01, Eccn)
