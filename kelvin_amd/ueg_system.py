"""Uniform electron gas in a plane-wave basis (integral provider).

Same basis ordering, integrals and generalised-orbital ('g') quantities as
kelvin/ueg_utils.py:21-138 and kelvin/ueg_system.py:31-131,266-267,400-419,
650-651, built with vectorised numpy instead of the reference's O(n^4)
Python loop (needed at n = 114 spin orbitals: 1.35 GB of <pq||rs>).
"""
import logging

import numpy

from . import ft_utils
from .system import System


class UEGBasis(object):
    """Plane waves 2*pi/L*(i,j,l) with k^2 < 2*cutoff, stable-sorted by |n|^2
    and truncated to norb (kelvin/ueg_utils.py:21-54)."""

    def __init__(self, L, cutoff, norb=None):
        self.L = L
        self.cutoff = cutoff
        kmax = numpy.sqrt(2 * cutoff)
        imax = int(numpy.ceil(L * kmax / (2 * numpy.pi)) + 0.1)
        r = numpy.arange(-imax, imax)
        i, j, l = numpy.meshgrid(r, r, r, indexing='ij')
        grid = numpy.stack((i.ravel(), j.ravel(), l.ravel()), axis=1)
        n2 = (grid * grid).sum(axis=1)
        k2 = 4.0 * numpy.pi * numpy.pi / (L * L) * n2
        grid, n2 = grid[k2 < 2.0 * cutoff], n2[k2 < 2.0 * cutoff]
        order = numpy.argsort(n2, kind='stable')
        basis = grid[order]
        if norb is not None:
            basis = basis[:norb]
        self.basis = [tuple(int(v) for v in x) for x in basis]
        self._ivec = basis.astype(numpy.int64)
        ccc = 2.0 * numpy.pi / L
        self.ks = [numpy.array((ccc * x[0], ccc * x[1], ccc * x[2]))
                   for x in self.basis]
        self.Es = [(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) / 2.0
                   for k in self.ks]

    def get_nbsf(self):
        return len(self.basis)

    def g_build_diag(self):
        e = numpy.asarray(self.Es)
        return numpy.hstack((e, e))

    def build_r2e_matrix(self):
        """Spatial <pq|rs> = 4 pi / (L^3 |k_p - k_r|^2) when
        k_p + k_q = k_r + k_s and p != r (kelvin/ueg_utils.py:83-101)."""
        nb = self.get_nbsf()
        X = self._ivec
        ks = numpy.asarray(self.ks)
        aaa = 4.0 * numpy.pi / (self.L * self.L * self.L)
        lo = X.min(axis=0)
        span = X.max(axis=0) - lo + 1
        table = -numpy.ones(tuple(span), dtype=numpy.int64)
        table[tuple((X - lo).T)] = numpy.arange(nb)
        V = numpy.zeros((nb, nb, nb, nb))
        p, q, r = numpy.meshgrid(numpy.arange(nb), numpy.arange(nb),
                                 numpy.arange(nb), indexing='ij')
        p, q, r = p.ravel(), q.ravel(), r.ravel()
        xs = X[p] + X[q] - X[r] - lo
        inside = numpy.all((xs >= 0) & (xs < span), axis=1) & (p != r)
        p, q, r, xs = p[inside], q[inside], r[inside], xs[inside]
        s = table[tuple(xs.T)]
        keep = s >= 0
        p, q, r, s = p[keep], q[keep], r[keep], s[keep]
        dk = ks[p] - ks[r]
        V[p, q, r, s] = aaa / (dk[:, 0] * dk[:, 0] + dk[:, 1] * dk[:, 1]
                               + dk[:, 2] * dk[:, 2])
        return V

    def build_g2e_matrix(self, anti=True):
        """Spin-orbital (alpha block, then beta) <pq||rs>
        (kelvin/ueg_utils.py:109-138)."""
        nb = self.get_nbsf()
        Vs = self.build_r2e_matrix()
        n = 2 * nb
        out = numpy.zeros((n, n, n, n))
        blk = [slice(0, nb), slice(nb, n)]
        VsT = Vs.transpose(0, 1, 3, 2) if anti else None
        for a in range(2):
            for b in range(2):
                out[blk[a], blk[b], blk[a], blk[b]] += Vs
                if anti:
                    out[blk[a], blk[b], blk[b], blk[a]] -= VsT
        return out


class UEGSystem(System):
    """The uniform electron gas (kelvin/ueg_system.py:14-65).

    Only the generalised-orbital quantities are provided; a system built
    with orbtype 'u' (the reference's default) is solved in the
    spin-orbital form, which gives the same grand potential for this
    spin-symmetric system (kelvin/tests/test_ft_ccsd.py:119-149).
    """

    def __init__(self, T, L, Emax, mu=None, na=None, nb=None, norb=None,
                 orbtype='u', madelung=None):
        self.T = T
        self.L = L
        self.basis = UEGBasis(L, Emax, norb=norb)
        if na is not None or nb is not None or not T > 0.0:
            # fixed occupations at T = 0 (kelvin/ueg_system.py:46-55) belong
            # to zero-temperature CCSD, which the GPU path does not build
            raise NotImplementedError(
                "UEGSystem needs T > 0 and mu (zero-temperature occupations "
                "are not part of the GPU path)")
        if mu is None:
            raise ValueError("UEGSystem needs the chemical potential mu")
        self.mu = mu
        N = ft_utils.ff(1.0 / self.T, self.g_energies_tot(), self.mu).sum()
        self.Na = N / 2.0
        self.Nb = self.Na
        self.N = self.Na + self.Nb
        self.den = self.N / (L * L * L)
        self.rs = (3 / (4.0 * numpy.pi * self.den)) ** (1.0 / 3.0)
        self.Ef = 0.5 * (3.0 * numpy.pi * numpy.pi * self.den) ** (2.0 / 3.0)
        self.Tf = self.T / self.Ef
        self.orbtype = orbtype
        self.madelung = madelung
        self._mconst = 2.837297479 / (2 * self.L)
        self._eri = None

    def has_g(self):
        return True

    def has_u(self):
        return self.orbtype != 'g'

    def has_r(self):
        return self.orbtype == 'r'

    def verify(self, T, mu):
        if T > 0.0:
            return T == self.T and mu == self.mu
        return T == self.T

    def const_energy(self):
        if self.madelung == 'const':
            return -(self.Na + self.Nb) * self._mconst
        return 0.0

    def g_energies_tot(self):
        return self.basis.g_build_diag()

    def g_aint_tot(self):
        if self._eri is None:
            self._eri = self.basis.build_g2e_matrix()
        return self._eri

    def g_int_tot(self):
        return self.basis.build_g2e_matrix(anti=False)

    def _fo(self):
        return ft_utils.ff(1.0 / self.T, self.g_energies_tot(), self.mu)

    def g_fock_tot(self):
        """Kinetic diagonal plus the finite-T exchange field
        sum_r <pr||qr> fo_r (kelvin/ueg_system.py:400-419)."""
        d = self.g_energies_tot()
        if self.T > 0.0:
            fo = self._fo()
        else:
            fo = (d < self.mu).astype(float)
        JK = numpy.einsum('prqr,r->pq', self.g_aint_tot(), fo)
        return numpy.diag(d) + JK

    def u_energies_tot(self):
        """(ea, eb): the plane-wave energies for each spin."""
        e = numpy.asarray(self.basis.Es)
        return e, e.copy()

    def u_fock_tot(self):
        """(fa, fb): the spin blocks of g_fock_tot (the Fock matrix is spin
        diagonal)."""
        f = self.g_fock_tot()
        nk = self.basis.get_nbsf()
        return f[:nk, :nk].copy(), f[nk:, nk:].copy()

    def u_aint_tot(self):
        """(Va, Vb, Vabab): antisymmetrised same-spin integrals and the
        alpha-beta <pq|rs> (whose exchange part vanishes by spin), the spin
        blocks of g_aint_tot (kelvin/ueg_system.py:266-267 builds the same
        from the spatial integrals)."""
        V = self.g_aint_tot()
        nk = self.basis.get_nbsf()
        a, b = slice(0, nk), slice(nk, 2 * nk)
        return (V[a, a, a, a].copy(), V[b, b, b, b].copy(),
                V[a, b, a, b].copy())

    def g_device_integrals(self, engine):
        """(fock, eri, E1) built on the GPU by the engine's library: eri in
        HBM, bitwise equal to g_aint_tot(); fock and E1 its mean field at
        (T, mu) as g_fock_tot() and get_mp1() define them (equal to their
        host values up to the order of the sums)."""
        b = self.basis
        eri = engine.ueg_eri(b._ivec, numpy.asarray(b.ks), self.L)
        d = self.g_energies_tot()
        fo = self._fo() if self.T > 0.0 else (d < self.mu).astype(float)
        fock, e1 = engine.mean_field(d, eri, fo)
        return fock, eri, e1

    def get_mp1(self):
        """1/2 sum_ij <ij||ij> fo_i fo_j (kelvin/ueg_system.py:126-131)."""
        if self.T > 0:
            fo = self._fo()
            return 0.5 * numpy.einsum('ijij,i,j->', self.g_aint_tot(), fo, fo)
        V = self.g_aint_tot()
        occ = self.g_energies_tot() < self.mu
        return 0.5 * numpy.einsum('ijij->', V[numpy.ix_(occ, occ, occ, occ)])

    def g_mp1_den(self):
        fo = self._fo()
        fv = ft_utils.ffv(1.0 / self.T, self.g_energies_tot(), self.mu)
        return -(1.0 / self.T) * numpy.einsum(
            'ijij,i,j->i', self.g_aint_tot(), fo * fv, fo)

    def g_fock_d_den(self):
        fo = self._fo()
        fv = ft_utils.ffv(1.0 / self.T, self.g_energies_tot(), self.mu)
        return numpy.einsum('piqi,i->pqi', self.g_aint_tot(), fo * fv)


class PUEGSystem(System):
    """The spin-polarised uniform electron gas (kelvin/pueg_system.py:10-203):
    one spin, so the generalised orbitals are the plane waves themselves,
    with energies k^2/2 (:114-115), <pq||rs> = <pq|rs> - <pq|sr> of the
    spatial integrals (:197-203), the Fock matrix diag(e) + sum_r <pr||qr>
    fo_r (:141-168), E1 = 1/2 sum_ij <ij||ij> fo_i fo_j (:71-80) and the
    derivative densities of E1 and of the Fock matrix (:95-103, :178-195).
    Finite temperature only, like UEGSystem."""

    def __init__(self, T, L, Emax, mu=None, n=None, norb=None):
        self.T = T
        self.L = L
        self.basis = UEGBasis(L, Emax, norb=norb)
        if n is not None or not T > 0.0:
            # a fixed electron count sets the zero-temperature occupations
            # (kelvin/pueg_system.py:37-41), which the GPU path does not build
            raise NotImplementedError(
                "PUEGSystem needs T > 0 and mu (zero-temperature occupations "
                "are not part of the GPU path)")
        if mu is None:
            raise ValueError("PUEGSystem needs the chemical potential mu")
        self.mu = mu
        self.N = ft_utils.ff(1.0 / self.T, self.g_energies_tot(),
                             self.mu).sum()
        self.den = self.N / (L * L * L)
        self.rs = (3 / (4.0 * numpy.pi * self.den)) ** (1.0 / 3.0)
        self.Ef = 0.5 * (3.0 * numpy.pi * numpy.pi * self.den) ** (2.0 / 3.0)
        self.Tf = self.T / self.Ef
        self.orbtype = 'g'
        self._eri = None

    def has_g(self):
        return True

    def has_u(self):
        return False

    def has_r(self):
        return False

    def verify(self, T, mu):
        if T > 0.0:
            return T == self.T and mu == self.mu
        return T == self.T

    def const_energy(self):
        return 0.0

    def g_energies_tot(self):
        return numpy.asarray(self.basis.Es)

    def g_aint_tot(self):
        if self._eri is None:
            V = self.basis.build_r2e_matrix()
            self._eri = V - V.transpose(0, 1, 3, 2)
        return self._eri

    def _fo(self):
        return ft_utils.ff(1.0 / self.T, self.g_energies_tot(), self.mu)

    def g_fock_tot(self):
        d = self.g_energies_tot()
        JK = numpy.einsum('prqr,r->pq', self.g_aint_tot(), self._fo())
        return numpy.diag(d) + JK

    def get_mp1(self):
        fo = self._fo()
        return 0.5 * numpy.einsum('ijij,i,j->', self.g_aint_tot(), fo, fo)

    def g_mp1_den(self):
        fo = self._fo()
        fv = ft_utils.ffv(1.0 / self.T, self.g_energies_tot(), self.mu)
        return -(1.0 / self.T) * numpy.einsum(
            'ijij,i,j->i', self.g_aint_tot(), fo * fv, fo)

    def g_fock_d_den(self):
        fo = self._fo()
        fv = ft_utils.ffv(1.0 / self.T, self.g_energies_tot(), self.mu)
        return numpy.einsum('piqi,i->pqi', self.g_aint_tot(), fo * fv)


def ueg_parameters(rs, N, t_over_tf):
    """(L, T, E_F) of an N-electron gas at Wigner-Seitz radius rs and
    reduced temperature T/T_F (L as kelvin/tests/test_ueg.py:8-10, E_F as
    kelvin/ueg_system.py:57-62)."""
    L = rs * (4.0 * numpy.pi * N / 3.0) ** (1.0 / 3.0)
    Ef = 0.5 * (3.0 * numpy.pi ** 2 * N / L ** 3) ** (2.0 / 3.0)
    return L, t_over_tf * Ef, Ef


def chemical_potential(en, beta, N, tol=1e-14):
    """mu with sum_p ff(beta, e_p, mu) = N (bisection; non-interacting
    occupations, the quantity UEGSystem derives N from)."""
    lo, hi = en.min() - 50.0 / beta, en.max() + 50.0 / beta
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if ft_utils.ff(beta, en, mid).sum() > N:
            hi = mid
        else:
            lo = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)


def ueg_from_config(nk, rs, N=14, t_over_tf=0.1, Emax=None, orbtype='g'):
    """UEGSystem with nk k-points (closed shells: 19, 27, 33, 57 ...) at
    (rs, T/T_F), mu solved for <N> = N."""
    L, T, Ef = ueg_parameters(rs, N, t_over_tf)
    if Emax is None:
        # the nk-th plane wave sits inside k^2 < 2 Emax with margin
        Emax = 0.5 * (2 * numpy.pi / L) ** 2 * 8.0
    basis = UEGBasis(L, Emax, norb=nk)
    if basis.get_nbsf() != nk:
        raise ValueError("cutoff too small for %d k-points" % nk)
    en = basis.g_build_diag()
    mu = chemical_potential(en, 1.0 / T, N)
    sys = UEGSystem(T, L, Emax, mu=mu, norb=nk, orbtype=orbtype)
    logging.info("UEG nk=%d rs=%.3f T/TF=%.3f: L=%.6f T=%.6f mu=%.6f N0=%.6f",
                 nk, rs, t_over_tf, L, T, mu, sys.N)
    return sys
