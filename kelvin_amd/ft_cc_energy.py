"""Correlation grand potential (kelvin/ft_cc_energy.py:7-32) on the GPU.

Omega_cc = (1/beta) sum_y g_y [ sum_ai T1[y,a,i] f[i,a]
            + sum_abij (T2[y]/4 + Qterm * T1[y,a,i] T1[y,b,j]/2) eri[i,j,a,b] ]
"""
import numpy
import torch

from .engine import FTEngine
from .ov_blocks import one_e_blocks, two_e_blocks


def _source(t):
    return getattr(t, "_ftcc_source", (None, None))


def ft_cc_energy(T1, T2, f, eri, g, beta, Qterm=True):
    ef, nf = _source(f)
    ee, ne = _source(eri)
    ng = T1.shape[0]
    if ef is not None and ef is ee and nf == "ov" and ne == "oovv":
        eng = ef
    else:
        # arbitrary F.ov / I.oovv: a context holding just these two blocks
        nv, no = T1.shape[1], T1.shape[2]
        eng = FTEngine(no, nv, ng)
        z = lambda *s: numpy.zeros(s)
        F = one_e_blocks(oo=z(no, no), ov=f, vo=z(nv, no), vv=z(nv, nv))
        I = two_e_blocks(vvvv=z(nv, nv, nv, nv), vvvo=z(nv, nv, nv, no),
                         vovv=z(nv, no, nv, nv), vvoo=z(nv, nv, no, no),
                         vovo=z(nv, no, nv, no), oovv=eri,
                         vooo=z(nv, no, no, no), ooov=z(no, no, no, nv),
                         oooo=z(no, no, no, no))
        eng.set_blocks(F, I)
        eng.set_grid(numpy.zeros(ng), numpy.zeros((ng, ng)))
    if eng.ng != ng:
        ti, G = getattr(eng, "_grid", (numpy.zeros(ng), numpy.zeros((ng, ng))))
        if len(ti) != ng:
            ti, G = numpy.zeros(ng), numpy.zeros((ng, ng))
        eng.set_grid(ti, G)
    return eng.energy(eng.to_device(T1), eng.to_device(T2), g, beta,
                      qterm=Qterm)


def ft_ucc_energy(T1a, T1b, T2aa, T2ab, T2bb, fa, fb, Ia, Ib, Iabab, g, beta,
                  Qterm=True):
    """Unrestricted form (kelvin/ft_cc_energy.py:35-72) from the spin blocks:
    fa/fb are the F.ov blocks, Ia/Ib/Iabab the oovv blocks.  Evaluated as
    the spin-orbital energy of the assembled amplitudes and integrals, which
    is the same sum block by block; with Qterm=False the reference weights
    the beta-beta integrals with T2aa (kelvin/ft_cc_energy.py:57), and so
    does this function."""
    import itertools
    from . import unrestricted
    noa, nva = fa.shape[0], fa.shape[1]
    nob, nvb = fb.shape[0], fb.shape[1]
    lay = unrestricted.SpinLayout(noa, nob, nva, nvb)
    if not Qterm:
        if tuple(T2aa.shape) != tuple(T2bb.shape):
            raise ValueError("Qterm=False needs T2aa and T2bb of one shape "
                             "(the beta-beta term uses T2aa)")
        T2bb = T2aa
    dev = torch.device("cuda", 0)
    T1, T2 = unrestricted.spin_orbital_amplitudes(lay, T1a, T1b, T2aa, T2ab,
                                                  T2bb, dev)
    f = torch.zeros(lay.size('o'), lay.size('v'), dtype=torch.float64,
                    device=dev)
    f[lay.sl('o', 0), lay.sl('v', 0)] = unrestricted._dev(fa, dev)
    f[lay.sl('o', 1), lay.sl('v', 1)] = unrestricted._dev(fb, dev)
    oovv = torch.zeros(lay.size('o'), lay.size('o'), lay.size('v'),
                       lay.size('v'), dtype=torch.float64, device=dev)
    ab = _Blocks(oovv=Iabab)
    for spins in itertools.product((0, 1), repeat=4):
        if spins[0] + spins[1] != spins[2] + spins[3]:
            continue
        idx = tuple(lay.sl(r, sp) for r, sp in zip("oovv", spins))
        if spins == (0, 0, 0, 0):
            oovv[idx] = unrestricted._dev(Ia, dev)
        elif spins == (1, 1, 1, 1):
            oovv[idx] = unrestricted._dev(Ib, dev)
        else:
            oovv[idx] = unrestricted._mixed_source("oovv", spins, ab, dev)
    return ft_cc_energy(T1, T2, f, oovv, g, beta, Qterm=Qterm)


class _Blocks(object):
    def __init__(self, **kw):
        self.__dict__.update(kw)
