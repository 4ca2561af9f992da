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
