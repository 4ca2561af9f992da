"""ftcc_energy_noalloc (the allocation-free entry point) against
ftcc_energy, through the C ABI, on two calls in a row on one GPU, the
second with a smaller problem; bit-exact (the same kernels and reduction
order)."""
import ctypes

import numpy
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_energy_noalloc_equals_energy():
    from kelvin_amd import _lib, quadrature
    from kelvin_amd.engine import FTEngine
    from kelvin_amd.ueg_system import ueg_from_config
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _lib.load()
    fn = lib.ftcc_energy_noalloc      # argument types from _lib.SIGNATURES
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    # 19 k-points (n = 38), ng = 6, then the smaller 7 k-points (n = 14), 4
    for nk, N, ng in ((19, 14, 6), (7, 6, 4)):
        sysg = ueg_from_config(nk, 2.0, N=N, t_over_tf=0.3)
        en = sysg.g_energies_tot()
        beta = 1.0 / sysg.T
        ti, g, G = quadrature.ft_quad(ng, beta, 'lin')
        eng = FTEngine(en.shape[0], en.shape[0], ng)
        fock, eri, _ = sysg.g_device_integrals(eng)
        eng.set_system(en, fock, eri, beta, sysg.mu)
        del fock, eri
        eng.set_grid(ti, G)
        T1, T2 = eng.empty_amplitudes()
        eng.driving_terms(T1, T2)
        eng.integrate(T1, T2)
        g_dev = torch.as_tensor(g, dtype=torch.float64, device=eng.device)
        out = torch.full((1,), numpy.nan, dtype=torch.float64,
                         device=eng.device)
        for qterm in (1, 0):
            ref = eng.energy(T1, T2, g, beta, qterm=bool(qterm))
            rc = fn(eng._h, vp(T1), vp(T2), vp(g_dev), float(beta), qterm,
                    vp(out))
            assert rc == 0, lib.ftcc_last_error(eng._h)
            torch.cuda.synchronize()
            got = float(out.cpu()[0])
            assert got == ref, (nk, qterm, got, ref)
            assert ref != 0.0
        eng.close()
