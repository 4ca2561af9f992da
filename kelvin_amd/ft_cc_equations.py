"""Imaginary-time CCSD amplitude equations on the GPU.

``ccsd_stanton`` mirrors kelvin/ft_cc_equations.py:96-113: Tbar = -F.vo -
R1(T), -I.vvoo - R2(T) at every grid point (Stanton-Gauss intermediates; the
reference delegates the residual to cqcpy cc_equations._Stanton), followed by
the exp(-tau D) integration of kelvin/quadrature.py:292-317.
"""
import torch

from . import unrestricted
from .engine import FTEngine
from .ov_blocks import EngineBlocks, EngineDenominators


def _shape_from(I):
    vvoo = I.vvoo
    nv, no = vvoo.shape[0], vvoo.shape[2]
    return no, nv


def engine_of(F, I, D1, D2, ng):
    """The FTEngine holding (F, I); user arrays are uploaded into a fresh
    engine.  Explicit D1/D2 arrays replace the on-the-fly denominators."""
    if isinstance(I, EngineBlocks):
        eng = I.engine
        if F is not None and isinstance(F, EngineBlocks) and F.engine is not eng:
            raise ValueError("F and I come from different engines")
    else:
        no, nv = _shape_from(I)
        eng = FTEngine(no, nv, ng)
        eng.set_blocks(F, I)
        if D1 is None or isinstance(D1, EngineDenominators):
            raise ValueError("explicit D1/D2 are required with user blocks")
    if D1 is not None and not isinstance(D1, EngineDenominators):
        eng.set_denominators(D1, D2)
    return eng


def ccsd_stanton(F, I, T1old, T2old, D1, D2, ti, ng, G):
    """New amplitudes (T1, T2) as device tensors (kelvin/ft_cc_equations.py:96)."""
    eng = engine_of(F, I, D1, D2, ng)
    eng.set_grid(ti, G)
    T1old = eng.to_device(T1old)
    T2old = eng.to_device(T2old)
    T1 = torch.empty_like(T1old)
    T2 = torch.empty_like(T2old)
    eng.residual(T1old, T2old, T1, T2)
    eng.integrate(T1, T2)
    return T1, T2


def ccsd_residual(F, I, T1, T2, D1=None, D2=None, ti=None, G=None):
    """Pre-integration residual (Tbar1, Tbar2) at every grid point."""
    ng = T1.shape[0]
    eng = engine_of(F, I, D1, D2, ng)
    if ti is not None:
        eng.set_grid(ti, G)
    elif eng.ng != ng:
        raise ValueError("pass ti/G: the engine grid has %d points" % eng.ng)
    T1 = eng.to_device(T1)
    T2 = eng.to_device(T2)
    T1b = torch.empty_like(T1)
    T2b = torch.empty_like(T2)
    eng.residual(T1, T2, T1b, T2b)
    return T1b, T2b


def unrestricted_engine(Fa, Fb, Ia, Ib, Iabab, D1a, D1b, D2aa, D2ab, D2bb, ng):
    """An engine holding the spin-orbital assembly of alpha/beta blocks, with
    spin-blocked products (labels: alpha orbitals first in each role)."""
    lay = unrestricted.layout_of(Fa, Fb)
    eng = FTEngine(lay.size('o'), lay.size('v'), ng)
    F, I = unrestricted.spin_orbital_blocks(lay, Fa, Fb, Ia, Ib, Iabab,
                                            eng.device)
    eng.set_blocks(F, I)
    del F, I
    D1, D2 = unrestricted.spin_orbital_denominators(lay, D1a, D1b, D2aa, D2ab,
                                                    D2bb, eng.device)
    eng.set_denominators(D1, D2)
    eng.set_spin(lay.labels('o'), lay.labels('v'))
    return eng, lay


def uccsd_stanton(Fa, Fb, Ia, Ib, Iabab, T1aold, T1bold, T2aaold, T2abold,
                  T2bbold, D1a, D1b, D2aa, D2ab, D2bb, ti, ng, G):
    """New unrestricted amplitudes ((T1a, T1b), (T2aa, T2ab, T2bb)) as device
    tensors (kelvin/ft_cc_equations.py:130-164)."""
    eng, lay = unrestricted_engine(Fa, Fb, Ia, Ib, Iabab, D1a, D1b, D2aa,
                                   D2ab, D2bb, ng)
    eng.set_grid(ti, G)
    T1, T2 = unrestricted.spin_orbital_amplitudes(
        lay, T1aold, T1bold, T2aaold, T2abold, T2bbold, eng.device)
    T1b = torch.empty_like(T1)
    T2b = torch.empty_like(T2)
    eng.residual(T1, T2, T1b, T2b)
    eng.integrate(T1b, T2b)
    return unrestricted.unrestricted_amplitudes(lay, T1b, T2b)
