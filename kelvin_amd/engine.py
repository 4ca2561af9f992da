"""FTEngine: one libftcc context bound to one GPU and torch's current stream.

Device memory for amplitudes is owned by torch tensors (float64, C order,
the reference's layouts T1[y,a,i], T2[y,a,b,i,j]); integrals and residual
workspace are owned by the library context.  Every method calls through the
C ABI of include/ftcc.h; nothing here computes on the host.
"""
import ctypes

import numpy
import torch

from . import _lib
from ._lib import FTCCError, PREC

BLOCK_INDEX = {"oo": 0, "ov": 1, "vo": 2, "vv": 3, "vvvv": 4, "vvvo": 5,
               "vovv": 6, "vvoo": 7, "vovo": 8, "oovv": 9, "vooo": 10,
               "ooov": 11, "oooo": 12}
TWO_E = ("vvvv", "vvvo", "vovv", "vvoo", "vovo", "oovv", "vooo", "ooov",
         "oooo")
ONE_E = ("oo", "ov", "vo", "vv")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _dbuf(x):
    a = numpy.ascontiguousarray(numpy.asarray(x, dtype=numpy.float64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_device():
    if not torch.cuda.is_available():
        raise FTCCError("kelvin_amd needs a ROCm GPU (torch.cuda is not "
                        "available); there is no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


def block_shape(name, no, nv):
    return tuple(no if ch == 'o' else nv for ch in name)


class FTEngine:
    """Finite-temperature CCSD amplitude engine for (no, nv, ng)."""

    def __init__(self, no, nv, ng, precision="fp64", device=None):
        if precision not in PREC:
            raise ValueError("precision must be one of %s" % sorted(PREC))
        self.no, self.nv, self.ng = int(no), int(nv), int(ng)
        self.precision = precision
        self.device = torch.device(device) if device is not None \
            else default_device()
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            rc = self._lib.ftcc_create(ctypes.byref(self._h), self.no,
                                       self.nv, self.ng, PREC[precision],
                                       ctypes.c_void_p(stream))
        if rc != 0:
            raise FTCCError("ftcc_create failed (code %d)" % rc)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ftcc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        with torch.cuda.device(self.device):
            rc = getattr(self._lib, name)(self._h, *args)
        _lib.check(rc, self._h, name)

    def _check(self, t, shape, name):
        if not isinstance(t, torch.Tensor):
            raise TypeError("%s must be a torch tensor on %s" % (name, self.device))
        if t.device != self.device or t.dtype != torch.float64:
            raise ValueError("%s must be float64 on %s" % (name, self.device))
        if tuple(t.shape) != tuple(shape) or not t.is_contiguous():
            raise ValueError("%s must be contiguous with shape %s, got %s"
                             % (name, tuple(shape), tuple(t.shape)))

    def t1_shape(self):
        return (self.ng, self.nv, self.no)

    def t2_shape(self):
        return (self.ng, self.nv, self.nv, self.no, self.no)

    def empty_amplitudes(self):
        z = dict(dtype=torch.float64, device=self.device)
        return torch.empty(self.t1_shape(), **z), torch.empty(self.t2_shape(), **z)

    def to_device(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(numpy.ascontiguousarray(x, dtype=numpy.float64),
                               device=self.device)

    # -- set-up --------------------------------------------------------------
    def set_precision(self, precision):
        self._call("ftcc_set_precision", PREC[precision])
        self.precision = precision

    def set_system(self, en, fock, eri, beta, mu, iocc=None, ivir=None):
        """Fermi-weighted integrals from (en, g_fock_tot, g_aint_tot);
        returns sum_p GP0(e_p)."""
        en = self.to_device(en)
        n = en.shape[0]
        fock = self.to_device(fock)
        eri = self.to_device(eri)
        self._check(fock, (n, n), "fock")
        self._check(eri, (n, n, n, n), "eri")
        io = numpy.arange(n, dtype=numpy.int32) if iocc is None \
            else numpy.asarray(iocc, dtype=numpy.int32)
        iv = numpy.arange(n, dtype=numpy.int32) if ivir is None \
            else numpy.asarray(ivir, dtype=numpy.int32)
        if io.size != self.no or iv.size != self.nv:
            raise ValueError("active lists do not match (no, nv)")
        om = ctypes.c_double()
        ip = ctypes.POINTER(ctypes.c_int)
        self._call("ftcc_set_system", n, _vp(en), _vp(fock), _vp(eri),
                   float(beta), float(mu), io.ctypes.data_as(ip),
                   iv.ctypes.data_as(ip), ctypes.byref(om))
        return om.value

    def set_blocks(self, F, I):
        """Arbitrary one_e_blocks / two_e_blocks (numpy or torch)."""
        no, nv = self.no, self.nv
        fs = []
        for name in ONE_E:
            t = self.to_device(getattr(F, name))
            self._check(t, block_shape(name, no, nv), "F." + name)
            fs.append(t)
        ts = []
        for name in TWO_E:
            t = self.to_device(getattr(I, name))
            self._check(t, block_shape(name, no, nv), "I." + name)
            ts.append(t)
        fp = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in fs])
        ip = (ctypes.c_void_p * 9)(*[t.data_ptr() for t in ts])
        self._call("ftcc_set_blocks", fp, ip)

    def set_spin(self, lab_o=None, lab_v=None):
        """Spin labels (0/1) of the occupied and virtual lists for blocked
        O(n^6) products; None for both switches blocking off.  Raises
        FTCCError if the integrals do not conserve spin under the labels."""
        if lab_o is None and lab_v is None:
            self._call("ftcc_set_spin", None, None)
            return
        lo = numpy.ascontiguousarray(lab_o, dtype=numpy.int32)
        lv = numpy.ascontiguousarray(lab_v, dtype=numpy.int32)
        if lo.shape != (self.no,) or lv.shape != (self.nv,):
            raise ValueError("spin labels need shapes (no,) and (nv,)")
        ip = ctypes.POINTER(ctypes.c_int)
        self._call("ftcc_set_spin", lo.ctypes.data_as(ip),
                   lv.ctypes.data_as(ip))

    def spin_blocked(self):
        """True when the O(n^6) products run as spin blocks."""
        return self.spin_mode() > 0

    def spin_mode(self):
        """0: dense; else bit flags: 1 O(n^6) products as spin blocks, 2 the
        O(n^5) contractions over spin-allowed blocks only, 4 ladder products
        over packed antisymmetric pairs, 8 a spin-flip symmetric system (for
        trusted amplitudes the ring products and the alpha-lead O(n^5) terms
        compute half their blocks, the rest being their flipped images)."""
        on = ctypes.c_int()
        self._call("ftcc_get_spin", ctypes.byref(on))
        return on.value

    def ueg_eri(self, ivec, kvec, L):
        """Antisymmetrised spin-orbital UEG integrals (alpha, then beta) of
        the plane waves ivec[nk,3] (integers), kvec[nk,3] built in HBM."""
        iv = numpy.ascontiguousarray(ivec, dtype=numpy.int32)
        kv = numpy.ascontiguousarray(kvec, dtype=numpy.float64)
        nk = iv.shape[0]
        if iv.shape != (nk, 3) or kv.shape != (nk, 3):
            raise ValueError("ivec and kvec need shape (nk, 3)")
        n = 2 * nk
        eri = torch.empty((n, n, n, n), dtype=torch.float64, device=self.device)
        self._call("ftcc_ueg_eri", nk,
                   iv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   kv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                   float(L), _vp(eri))
        return eri

    def mean_field(self, en, eri, fo):
        """(fock, E1): diag(en) + sum_r <pr||qr> fo_r and
        1/2 sum_ij <ij||ij> fo_i fo_j from device integrals."""
        en = self.to_device(en)
        fo = self.to_device(fo)
        n = en.shape[0]
        self._check(eri, (n, n, n, n), "eri")
        fock = torch.empty((n, n), dtype=torch.float64, device=self.device)
        e1 = ctypes.c_double()
        self._call("ftcc_mean_field", n, _vp(en), _vp(eri), _vp(fo),
                   _vp(fock), ctypes.byref(e1))
        return fock, e1.value

    def set_denominators(self, D1, D2):
        D1 = self.to_device(D1)
        D2 = self.to_device(D2)
        self._check(D1, (self.nv, self.no), "D1")
        self._check(D2, (self.nv, self.nv, self.no, self.no), "D2")
        self._call("ftcc_set_denominators", _vp(D1), _vp(D2))

    def set_grid(self, ti, G):
        """Set (ti, G); the grid size may change (amplitudes then have
        len(ti) slices)."""
        ti_a, ti_p = _dbuf(ti)
        G_a, G_p = _dbuf(G)
        ng = ti_a.shape[0]
        if ti_a.ndim != 1 or G_a.shape != (ng, ng):
            raise ValueError("grid shapes must be (ng,) and (ng, ng)")
        self._call("ftcc_set_grid", ng, ti_p, G_p)
        self.ng = ng
        self._grid = (ti_a.copy(), G_a.copy())

    def get_block(self, name):
        shape = block_shape(name, self.no, self.nv)
        out = torch.empty(shape, dtype=torch.float64, device=self.device)
        self._call("ftcc_get_block", BLOCK_INDEX[name], _vp(out))
        return out

    # -- hot path --------------------------------------------------------------
    def driving_terms(self, T1b, T2b):
        self._check(T1b, self.t1_shape(), "T1bar")
        self._check(T2b, self.t2_shape(), "T2bar")
        self._call("ftcc_driving_terms", _vp(T1b), _vp(T2b))

    def residual(self, T1, T2, T1b, T2b, skip_mask=0):
        for t, s, nm in ((T1, self.t1_shape(), "T1"), (T2, self.t2_shape(), "T2"),
                         (T1b, self.t1_shape(), "T1bar"),
                         (T2b, self.t2_shape(), "T2bar")):
            self._check(t, s, nm)
        self._call("ftcc_residual", _vp(T1), _vp(T2), _vp(T1b), _vp(T2b),
                   ctypes.c_uint64(skip_mask))

    def integrate(self, T1=None, T2=None):
        if T1 is not None:
            self._check(T1, self.t1_shape(), "T1bar")
        if T2 is not None:
            self._check(T2, self.t2_shape(), "T2bar")
        self._call("ftcc_integrate",
                   _vp(T1) if T1 is not None else ctypes.c_void_p(),
                   _vp(T2) if T2 is not None else ctypes.c_void_p())

    def energy(self, T1, T2, g, beta, qterm=True):
        self._check(T1, self.t1_shape(), "T1")
        self._check(T2, self.t2_shape(), "T2")
        g_a, g_p = _dbuf(g)
        out = ctypes.c_double()
        self._call("ftcc_energy", _vp(T1), _vp(T2), g_p, float(beta),
                   int(bool(qterm)), ctypes.byref(out))
        return out.value

    def sumsq(self, x):
        if x.device != self.device or x.dtype != torch.float64 \
                or not x.is_contiguous():
            raise ValueError("sumsq needs a contiguous float64 device tensor")
        out = ctypes.c_double()
        self._call("ftcc_sumsq", _vp(x), ctypes.c_int64(x.numel()),
                   ctypes.byref(out))
        return out.value

    def step(self, T1old, T2old, T1b, T2b, g, beta, damp, skip_mask=0):
        """One damped fixed-point body; returns (d1, d2, n1, n2, E)."""
        for t, s, nm in ((T1old, self.t1_shape(), "T1old"),
                         (T2old, self.t2_shape(), "T2old"),
                         (T1b, self.t1_shape(), "T1bar"),
                         (T2b, self.t2_shape(), "T2bar")):
            self._check(t, s, nm)
        g_a, g_p = _dbuf(g)
        out = (ctypes.c_double * 5)()
        self._call("ftcc_step", _vp(T1old), _vp(T2old), _vp(T1b), _vp(T2b),
                   g_p, float(beta), float(damp), ctypes.c_uint64(skip_mask),
                   out)
        return tuple(out)

    def step_graph(self, T1old, T2old, T1b, T2b, beta, damp, skip_mask=0):
        """``step`` replayed as a HIP graph (weights from ``set_weights``;
        the engine must live on a non-default stream; spin layout needs
        ``set_trusted(True)``)."""
        for t, s, nm in ((T1old, self.t1_shape(), "T1old"),
                         (T2old, self.t2_shape(), "T2old"),
                         (T1b, self.t1_shape(), "T1bar"),
                         (T2b, self.t2_shape(), "T2bar")):
            self._check(t, s, nm)
        out = (ctypes.c_double * 5)()
        self._call("ftcc_step_graph", _vp(T1old), _vp(T2old), _vp(T1b),
                   _vp(T2b), float(beta), float(damp),
                   ctypes.c_uint64(skip_mask), out)
        return tuple(out)

    def set_weights(self, g):
        """Quadrature weights for ``step_graph``."""
        g_a, g_p = _dbuf(g)
        self._call("ftcc_set_weights", g_p)

    def set_trusted(self, on=True):
        """Promise spin-conserving amplitudes: no counting pass, no host
        synchronisation in the residual."""
        self._call("ftcc_set_trusted", int(bool(on)))

    def set_correction(self, dT1b=None, dT2b=None):
        """Add fixed corrections to the residual of every later ``step``
        (None, None clears them); the tensors must stay alive meanwhile."""
        if (dT1b is None) != (dT2b is None):
            raise ValueError("give both corrections or neither")
        if dT1b is None:
            self._corr = None
            self._call("ftcc_set_correction", None, None)
            return
        self._check(dT1b, self.t1_shape(), "dT1bar")
        self._check(dT2b, self.t2_shape(), "dT2bar")
        self._corr = (dT1b, dT2b)
        self._call("ftcc_set_correction", _vp(dT1b), _vp(dT2b))

    def defect(self, T1, T2, low, dT1b, dT2b, S1, S2, skip_mask=0):
        """dTbar = Tbar_fp64(T) - Tbar_low(T) at every grid point (S1/S2
        amplitude-size scratch); the engine's precision is unchanged."""
        self._check_amps(T1, T2, dT1b, dT2b)
        self._check(S1, self.t1_shape(), "S1")
        self._check(S2, self.t2_shape(), "S2")
        self._call("ftcc_defect", _vp(T1), _vp(T2), PREC[low], _vp(dT1b),
                   _vp(dT2b), _vp(S1), _vp(S2), ctypes.c_uint64(skip_mask))

    # -- adjoint ---------------------------------------------------------------
    def adjoint_begin(self, T1, T2):
        """Fix the amplitudes of the adjoint (kept by reference: unchanged
        until ``adjoint_end``) and build their intermediates."""
        self._check(T1, self.t1_shape(), "T1")
        self._check(T2, self.t2_shape(), "T2")
        self._adj_T = (T1, T2)
        self._call("ftcc_adjoint_begin", _vp(T1), _vp(T2))

    def adjoint_cached(self):
        """Grid points whose adjoint intermediates are kept."""
        n = ctypes.c_int()
        self._call("ftcc_adjoint_cached", ctypes.byref(n))
        return n.value

    def adjoint_apply(self, Lt1, Lt2, G1, G2):
        """(G1, G2) = J^T (Lt1, Lt2) at every grid point (Lt2 antisymmetric;
        G2 projected onto antisymmetric tensors)."""
        self._check_amps(Lt1, Lt2, G1, G2)
        self._call("ftcc_adjoint_apply", _vp(Lt1), _vp(Lt2), _vp(G1),
                   _vp(G2))

    def adjoint_end(self):
        self._call("ftcc_adjoint_end")
        self._adj_T = None

    def adjoint_precision(self, precision):
        """Precision of the large blocks of adjoint_apply: "fp64" or
        "bf16x3"."""
        self._call("ftcc_adjoint_precision", PREC[precision])

    def integrate_adjoint(self, L1, L2, out1, out2):
        """out[x] = sum_y G[y,x] exp(-(t_y - t_x) D) L[y] (x < y; weight 1
        otherwise): the transpose of ``integrate``."""
        self._check_amps(L1, L2, out1, out2)
        self._call("ftcc_integrate_adjoint", _vp(L1), _vp(L2), _vp(out1),
                   _vp(out2))

    def density1(self, T1, T2, Lt1, Lt2, g, beta):
        """The unrelaxed one-particle response densities (pia [no,nv],
        pba [nv,nv], pji [no,no], pai [nv,no], kelvin's ccsd_1rdm) from the
        amplitudes and the integrated adjoint amplitudes Lt =
        integrate_adjoint(Lam); device tensors."""
        self._check_amps(T1, T2, Lt1, Lt2)
        no, nv = self.no, self.nv
        dv = dict(dtype=torch.float64, device=self.device)
        pia = torch.empty((no, nv), **dv)
        pba = torch.empty((nv, nv), **dv)
        pji = torch.empty((no, no), **dv)
        pai = torch.empty((nv, no), **dv)
        g_a, g_p = _dbuf(g)
        self._call("ftcc_density1", _vp(T1), _vp(T2), _vp(Lt1), _vp(Lt2),
                   g_p, float(beta), _vp(pia), _vp(pba), _vp(pji), _vp(pai))
        return pia, pba, pji, pai

    def lambda_update(self, L1, L2, G1, G2, T1, g, beta, damp):
        """L <- damp L + (1 - damp) (dE/dT - G); returns the squared norms
        (|dL1|^2, |dL2|^2, |L1|^2, |L2|^2) in kelvin's lambda scale."""
        self._check_amps(L1, L2, G1, G2)
        self._check(T1, self.t1_shape(), "T1")
        g_a, g_p = _dbuf(g)
        out = (ctypes.c_double * 4)()
        self._call("ftcc_lambda_update", _vp(L1), _vp(L2), _vp(G1), _vp(G2),
                   _vp(T1), g_p, float(beta), float(damp), out)
        return tuple(out)

    def lambda_guess(self, T1, g, beta, L1, L2):
        """kelvin's first lambda amplitudes (kelvin/ft_cc_equations.py:
        502-512) in the adjoint layout: -(1/beta) dE/dT."""
        self._check_amps(T1, L2, L1, L2)
        g_a, g_p = _dbuf(g)
        self._call("ftcc_lambda_guess", _vp(T1), g_p, float(beta), _vp(L1),
                   _vp(L2))

    def lambda_to_kelvin(self, L1, L2, g, beta):
        """kelvin's L1[y,i,a], L2[y,i,j,a,b] (scale of kelvin/ccsd.py:
        1113-1146) from the adjoint amplitudes."""
        self._check(L1, self.t1_shape(), "Lam1")
        self._check(L2, self.t2_shape(), "Lam2")
        z = dict(dtype=torch.float64, device=self.device)
        K1 = torch.empty((self.ng, self.no, self.nv), **z)
        K2 = torch.empty((self.ng, self.no, self.no, self.nv, self.nv), **z)
        g_a, g_p = _dbuf(g)
        self._call("ftcc_lambda_convert", 1, g_p, float(beta), _vp(L1),
                   _vp(L2), _vp(K1), _vp(K2))
        return K1, K2

    def lambda_from_kelvin(self, K1, K2, g, beta):
        """The adjoint amplitudes of kelvin's L1, L2."""
        K1 = self.to_device(K1)
        K2 = self.to_device(K2)
        self._check(K1, (self.ng, self.no, self.nv), "L1")
        self._check(K2, (self.ng, self.no, self.no, self.nv, self.nv), "L2")
        L1, L2 = self.empty_amplitudes()
        g_a, g_p = _dbuf(g)
        self._call("ftcc_lambda_convert", 0, g_p, float(beta), _vp(K1),
                   _vp(K2), _vp(L1), _vp(L2))
        return L1, L2

    def grid_derivative(self, L1, L2, T1b, T2b, Gd, beta):
        """sum_yx L_y (Gd[y,x] - G[y,x] (t_y - t_x) D / beta) P_yx Tbar_x."""
        self._check_amps(L1, L2, T1b, T2b)
        Gd_a, Gd_p = _dbuf(Gd)
        if Gd_a.shape != (self.ng, self.ng):
            raise ValueError("Gd must be (ng, ng)")
        out = ctypes.c_double()
        self._call("ftcc_grid_derivative", _vp(L1), _vp(L2), _vp(T1b),
                   _vp(T2b), Gd_p, float(beta), ctypes.byref(out))
        return out.value

    def dot(self, x, y):
        if x.shape != y.shape or not (x.is_contiguous() and y.is_contiguous()):
            raise ValueError("dot needs two contiguous tensors of one shape")
        for t in (x, y):
            if t.device != self.device or t.dtype != torch.float64:
                raise ValueError("dot needs float64 device tensors")
        out = ctypes.c_double()
        self._call("ftcc_dot", _vp(x), _vp(y), ctypes.c_int64(x.numel()),
                   ctypes.byref(out))
        return out.value

    def set_system_response(self, en, fock, dfock, eri, dx, beta, mu,
                            iocc=None, ivir=None):
        """Replace the integrals by their derivative along a change of
        x_p = beta (e_p - mu) by dx_p, with the bare Fock matrix moving by
        dfock (a later ``set_system`` restores them)."""
        en = self.to_device(en)
        n = en.shape[0]
        fock = self.to_device(fock)
        dfock = self.to_device(dfock)
        dx = self.to_device(dx)
        self._check(fock, (n, n), "fock")
        self._check(dfock, (n, n), "dfock")
        self._check(eri, (n, n, n, n), "eri")
        self._check(dx, (n,), "dx")
        io = numpy.arange(n, dtype=numpy.int32) if iocc is None \
            else numpy.asarray(iocc, dtype=numpy.int32)
        iv = numpy.arange(n, dtype=numpy.int32) if ivir is None \
            else numpy.asarray(ivir, dtype=numpy.int32)
        if io.size != self.no or iv.size != self.nv:
            raise ValueError("active lists do not match (no, nv)")
        ip = ctypes.POINTER(ctypes.c_int)
        self._call("ftcc_set_system_response", n, _vp(en), _vp(fock),
                   _vp(dfock), _vp(eri), _vp(dx), float(beta), float(mu),
                   io.ctypes.data_as(ip), iv.ctypes.data_as(ip))

    def set_chunk(self, nb):
        """Grid points per residual chunk (0: automatic)."""
        self._call("ftcc_set_chunk", int(nb))

    def chunk(self):
        """Grid points per residual chunk in use."""
        nb = ctypes.c_int()
        self._call("ftcc_get_chunk", ctypes.byref(nb))
        return nb.value

    def _check_amps(self, T1, T2, T1b, T2b):
        for t, s, nm in ((T1, self.t1_shape(), "T1"),
                         (T2, self.t2_shape(), "T2"),
                         (T1b, self.t1_shape(), "T1bar"),
                         (T2b, self.t2_shape(), "T2bar")):
            self._check(t, s, nm)

    def point_begin(self, ig, T1, T2, T1b, T2b):
        """Prepare grid point ig of the pointwise schedule (driving terms,
        first-point integral or linear extrapolation)."""
        self._check_amps(T1, T2, T1b, T2b)
        self._call("ftcc_point_begin", int(ig), _vp(T1), _vp(T2), _vp(T1b),
                   _vp(T2b))

    def point_step(self, ig, T1, T2, T1b, T2b, damp):
        """One iteration at grid point ig; returns (|dT1|^2, |dT2|^2)."""
        self._check_amps(T1, T2, T1b, T2b)
        out = (ctypes.c_double * 2)()
        self._call("ftcc_point_step", int(ig), _vp(T1), _vp(T2), _vp(T1b),
                   _vp(T2b), float(damp), out)
        return out[0], out[1]

    def td_rhs(self, T1, T2, K1, K2, singles=True):
        """K = Tbar(T) - D T at every slice (TD-CCSD right-hand side)."""
        self._check_amps(T1, T2, K1, K2)
        self._call("ftcc_td_rhs", _vp(T1), _vp(T2), _vp(K1), _vp(K2),
                   int(bool(singles)))

    def td_step(self, T1, T2, h, order=4, singles=True):
        """One explicit Runge-Kutta step (order 1, 2 or 4) of size h, in
        place, at every slice."""
        self._check(T1, self.t1_shape(), "T1")
        self._check(T2, self.t2_shape(), "T2")
        self._call("ftcc_td_step", int(order), float(h), int(bool(singles)),
                   _vp(T1), _vp(T2))

    def td_implicit_step(self, T1, T2, h, scheme, max_iter, damp, thresh,
                         singles=True):
        """One implicit step ('be', 'cn' or 'am2') of size h, in place;
        returns the number of inner iterations."""
        self._check(T1, self.t1_shape(), "T1")
        self._check(T2, self.t2_shape(), "T2")
        code = {"be": 0, "cn": 1, "am2": 2}[scheme]
        it = ctypes.c_int()
        self._call("ftcc_td_implicit_step", code, float(h), int(max_iter),
                   float(damp), float(thresh), int(bool(singles)), _vp(T1),
                   _vp(T2), ctypes.byref(it))
        return it.value

    def profile(self, on=True):
        """Bracket every O(n^6) product launch with HIP events."""
        self._call("ftcc_profile_enable", int(bool(on)))

    def profile_read(self):
        """(launches, device ms, algorithmic flops) since the last read."""
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        fl = ctypes.c_double()
        self._call("ftcc_profile_read", ctypes.byref(n), ctypes.byref(ms),
                   ctypes.byref(fl))
        return n.value, ms.value, fl.value

    def set_gemm_plan(self, tile_policy=0, split_target=1024):
        """Launch plan of the generic GEMMs (see ``ftcc_set_gemm_plan``):
        the knobs the tile autotuner of the batch searches."""
        self._call("ftcc_set_gemm_plan", int(tile_policy), int(split_target))

    PROFILE_CLASSES = ("products", "n5_gemms", "gathers", "assembly",
                       "update")

    def profile_classes(self):
        """{class: (launches, device ms, algorithmic flops, algorithmic
        bytes)} since the last read, for every kernel class."""
        out = {}
        for k, name in enumerate(self.PROFILE_CLASSES):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            fl = ctypes.c_double()
            by = ctypes.c_double()
            self._call("ftcc_profile_class", k, ctypes.byref(n),
                       ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by))
            out[name] = (n.value, ms.value, fl.value, by.value)
        return out

    def workspace_bytes(self):
        b = ctypes.c_int64()
        self._call("ftcc_workspace_bytes", ctypes.byref(b))
        return b.value

    def gemm(self, A, B, C, alpha=1.0, beta=0.0, precision=None):
        """C[b] = alpha A[b] @ B[b] + beta C[b] for 2-D or 3-D (batched)
        tensors with arbitrary strides (testing hook for the MFMA engine)."""
        def st(t):
            if t.dim() == 2:
                return 1, 0, t.stride(0), t.stride(1)
            return t.shape[0], t.stride(0), t.stride(1), t.stride(2)
        for t in (A, B, C):
            if t.device != self.device or t.dtype != torch.float64:
                raise ValueError("gemm operands must be float64 on device")
        nb, sAb, sAm, sAk = st(A)
        _, sBb, sBk, sBn = st(B)
        _, sCb, sCm, sCn = st(C)
        M, K = A.shape[-2], A.shape[-1]
        N = B.shape[-1]
        prec = PREC[precision or self.precision]
        self._call("ftcc_gemm", M, N, K, nb, float(alpha), _vp(A), sAm, sAk,
                   sAb, _vp(B), sBk, sBn, sBb, float(beta), _vp(C), sCm, sCn,
                   sCb, prec)
        return C
