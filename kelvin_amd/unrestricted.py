"""Unrestricted (alpha/beta) interface of the amplitude engine.

kelvin's 'u' path (kelvin/ft_cc_equations.py:130-164 uccsd_stanton,
kelvin/cc_utils.py:245-317 ft_ucc_iter, :696-777 uft_integrals,
kelvin/ft_cc_energy.py:35-72 ft_ucc_energy) keeps the amplitudes and
integrals as spin blocks: T1a, T1b, T2aa, T2ab, T2bb; Ia, Ib (antisymmetrised
same-spin integrals) and Iabab (the alpha-beta integrals <ab|cd>, 16 o/v
block types).  The same equations in spin orbitals -- alpha orbitals first,
then beta -- are what the engine solves, with its O(n^6) products run as
spin blocks, so this module only assembles spin-orbital operands from the
blocks and takes the blocks back out.  Every mixed-spin block of a
spin-orbital tensor is an (anti)symmetry image of an alpha-beta block: a
pair whose spins read (beta, alpha) is swapped, with a sign for each swap of
an antisymmetric tensor.
"""
import itertools

import torch

from .ov_blocks import one_e_blocks, two_e_blocks


class two_e_blocks_full(object):
    """The 16 o/v blocks of the alpha-beta integrals (cqcpy
    two_e_blocks_full, kelvin/cc_utils.py:769-775)."""

    NAMES = ("vvvv", "vvvo", "vvov", "vovv", "ovvv", "vvoo", "vovo", "ovvo",
             "voov", "ovov", "oovv", "vooo", "ovoo", "oovo", "ooov", "oooo")

    def __init__(self, **blocks):
        for k in self.NAMES:
            setattr(self, k, blocks.get(k))


def _dev(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64)
    return torch.as_tensor(x, dtype=torch.float64, device=device)


class SpinLayout(object):
    """Orbital counts per spin in the occupied and virtual roles; the
    spin-orbital lists are [alpha..., beta...] in each role."""

    def __init__(self, noa, nob, nva, nvb):
        self.no = {0: noa, 1: nob}
        self.nv = {0: nva, 1: nvb}

    def count(self, role, spin):
        return self.no[spin] if role == 'o' else self.nv[spin]

    def sl(self, role, spin):
        lo = 0 if spin == 0 else self.count(role, 0)
        return slice(lo, lo + self.count(role, spin))

    def size(self, role):
        return self.count(role, 0) + self.count(role, 1)

    def labels(self, role):
        return [0] * self.count(role, 0) + [1] * self.count(role, 1)


def _mixed_source(name, spins, Iab, device):
    """Block `name` (4 o/v letters) of the spin-orbital integrals for the
    mixed spin pattern `spins` as an image of the alpha-beta blocks."""
    letters, perm, sign = list(name), [0, 1, 2, 3], 1.0
    s = list(spins)
    if s[0] == 1:                      # (beta, alpha): swap the first pair
        letters[0], letters[1] = letters[1], letters[0]
        perm[0], perm[1] = perm[1], perm[0]
        sign = -sign
    if s[2] == 1:                      # swap the second pair
        letters[2], letters[3] = letters[3], letters[2]
        perm[2], perm[3] = perm[3], perm[2]
        sign = -sign
    src = _dev(getattr(Iab, "".join(letters)), device)
    # element [p0,p1,p2,p3] of the target is src at the swapped positions
    inv = [perm.index(k) for k in range(4)]
    return sign * src.permute(*inv)


def spin_orbital_blocks(lay, Fa, Fb, Ia, Ib, Iabab, device):
    """(F, I) spin-orbital one_e_blocks / two_e_blocks on `device`."""
    F = {}
    for name in ("oo", "ov", "vo", "vv"):
        r0, r1 = name
        t = torch.zeros(lay.size(r0), lay.size(r1), dtype=torch.float64,
                        device=device)
        for spin, Fx in ((0, Fa), (1, Fb)):
            t[lay.sl(r0, spin), lay.sl(r1, spin)] = _dev(getattr(Fx, name),
                                                         device)
        F[name] = t
    I = {}
    for name in ("vvvv", "vvvo", "vovv", "vvoo", "vovo", "oovv", "vooo",
                 "ooov", "oooo"):
        t = torch.zeros(*[lay.size(r) for r in name], dtype=torch.float64,
                        device=device)
        for spins in itertools.product((0, 1), repeat=4):
            if spins[0] + spins[1] != spins[2] + spins[3]:
                continue
            idx = tuple(lay.sl(r, sp) for r, sp in zip(name, spins))
            if spins == (0, 0, 0, 0):
                src = _dev(getattr(Ia, name), device)
            elif spins == (1, 1, 1, 1):
                src = _dev(getattr(Ib, name), device)
            else:
                src = _mixed_source(name, spins, Iabab, device)
            t[idx] = src
        I[name] = t
    return one_e_blocks(**F), two_e_blocks(**I)


def spin_orbital_amplitudes(lay, T1a, T1b, T2aa, T2ab, T2bb, device):
    """(T1, T2) in spin orbitals, leading grid axis kept."""
    T1a, T1b = _dev(T1a, device), _dev(T1b, device)
    T2aa, T2ab, T2bb = (_dev(x, device) for x in (T2aa, T2ab, T2bb))
    ng = T1a.shape[0]
    nv, no = lay.size('v'), lay.size('o')
    T1 = torch.zeros(ng, nv, no, dtype=torch.float64, device=device)
    T1[:, lay.sl('v', 0), lay.sl('o', 0)] = T1a
    T1[:, lay.sl('v', 1), lay.sl('o', 1)] = T1b
    T2 = torch.zeros(ng, nv, nv, no, no, dtype=torch.float64, device=device)
    va, vb = lay.sl('v', 0), lay.sl('v', 1)
    oa, ob = lay.sl('o', 0), lay.sl('o', 1)
    T2[:, va, va, oa, oa] = T2aa
    T2[:, vb, vb, ob, ob] = T2bb
    T2[:, va, vb, oa, ob] = T2ab
    T2[:, va, vb, ob, oa] = -T2ab.permute(0, 1, 2, 4, 3)
    T2[:, vb, va, oa, ob] = -T2ab.permute(0, 2, 1, 3, 4)
    T2[:, vb, va, ob, oa] = T2ab.permute(0, 2, 1, 4, 3)
    return T1, T2


def unrestricted_amplitudes(lay, T1, T2):
    """((T1a, T1b), (T2aa, T2ab, T2bb)) taken out of spin-orbital ones."""
    va, vb = lay.sl('v', 0), lay.sl('v', 1)
    oa, ob = lay.sl('o', 0), lay.sl('o', 1)
    return ((T1[:, va, oa].contiguous(), T1[:, vb, ob].contiguous()),
            (T2[:, va, va, oa, oa].contiguous(),
             T2[:, va, vb, oa, ob].contiguous(),
             T2[:, vb, vb, ob, ob].contiguous()))


def spin_orbital_denominators(lay, D1a, D1b, D2aa, D2ab, D2bb, device):
    """D1 [nv,no] and D2 [nv,nv,no,no] in spin orbitals (D2 is symmetric
    under swaps within a pair, so the mixed blocks carry no sign)."""
    nv, no = lay.size('v'), lay.size('o')
    va, vb = lay.sl('v', 0), lay.sl('v', 1)
    oa, ob = lay.sl('o', 0), lay.sl('o', 1)
    D1 = torch.zeros(nv, no, dtype=torch.float64, device=device)
    D1[va, oa] = _dev(D1a, device)
    D1[vb, ob] = _dev(D1b, device)
    D2 = torch.zeros(nv, nv, no, no, dtype=torch.float64, device=device)
    d2ab = _dev(D2ab, device)
    D2[va, va, oa, oa] = _dev(D2aa, device)
    D2[vb, vb, ob, ob] = _dev(D2bb, device)
    D2[va, vb, oa, ob] = d2ab
    D2[va, vb, ob, oa] = d2ab.permute(0, 1, 3, 2)
    D2[vb, va, oa, ob] = d2ab.permute(1, 0, 2, 3)
    D2[vb, va, ob, oa] = d2ab.permute(1, 0, 3, 2)
    return D1, D2


def layout_of(Fa, Fb):
    nva, noa = Fa.vo.shape
    nvb, nob = Fb.vo.shape
    return SpinLayout(noa, nob, nva, nvb)


_STORED = ("vvvv", "vvvo", "vovv", "vvoo", "vovo", "oovv", "vooo", "ooov",
           "oooo")


def alpha_beta_block(I, lay, name):
    """Block `name` of the alpha-beta integrals <ab|cd> (spins alpha, beta,
    alpha, beta) out of spin-orbital blocks I, through the swaps within a
    pair that lead to a stored block type."""
    for s1, s2 in ((0, 0), (1, 0), (0, 1), (1, 1)):
        letters, spins, perm = list(name), [0, 1, 0, 1], [0, 1, 2, 3]
        if s1:
            letters[0], letters[1] = letters[1], letters[0]
            spins[0], spins[1] = spins[1], spins[0]
            perm[0], perm[1] = perm[1], perm[0]
        if s2:
            letters[2], letters[3] = letters[3], letters[2]
            spins[2], spins[3] = spins[3], spins[2]
            perm[2], perm[3] = perm[3], perm[2]
        key = "".join(letters)
        if key in _STORED:
            src = getattr(I, key)
            idx = tuple(lay.sl(r, sp) for r, sp in zip(letters, spins))
            blk = src[idx].permute(*[perm.index(k) for k in range(4)])
            sign = -1.0 if (s1 + s2) % 2 else 1.0
            return (sign * blk).contiguous()
    raise KeyError(name)


def unrestricted_blocks(lay, F, I):
    """(Fa, Fb, Ia, Ib, Iabab) out of spin-orbital blocks (F, I)."""
    out_F, out_I = [], []
    for spin in (0, 1):
        out_F.append(one_e_blocks(**{
            k: getattr(F, k)[lay.sl(k[0], spin), lay.sl(k[1], spin)]
            .contiguous() for k in ("oo", "ov", "vo", "vv")}))
        out_I.append(two_e_blocks(**{
            k: getattr(I, k)[tuple(lay.sl(r, spin) for r in k)].contiguous()
            for k in _STORED}))
    Iabab = two_e_blocks_full(**{k: alpha_beta_block(I, lay, k)
                                 for k in two_e_blocks_full.NAMES})
    return out_F[0], out_F[1], out_I[0], out_I[1], Iabab
