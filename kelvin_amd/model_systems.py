"""Synthetic molecule-like provider with density-fitted integrals.

norb spatial orbitals: energies ~ U[-1, 1] (sorted) drawn first from
numpy.random.default_rng(seed), then B[Q,p,q] (Q < 2*norb) with the upper
triangle ~ N(0, sigma^2) mirrored to be symmetric in pq.  Chemists' integrals
(pq|rs) = sum_Q B_Qpq B_Qrs are spin-blocked (alpha block, then beta) and
antisymmetrised, <pq||rs> = (pr|qs) d(s_p,s_r) d(s_q,s_s) - (ps|qr) ...
The Fock matrix is diag(energies) (so the Fermi-weighted Fock blocks vanish)
and E1 = 1/2 sum_ij <ij||ij> fo_i fo_j.
"""
import numpy

from . import ft_utils
from .system import System


class DFModelSystem(System):
    def __init__(self, T, mu, norb=8, seed=0, sigma=0.03):
        rng = numpy.random.default_rng(seed)
        self.T, self.mu = T, mu
        e = numpy.sort(rng.uniform(-1.0, 1.0, norb))
        B = numpy.zeros((2 * norb, norb, norb))
        iu = numpy.triu_indices(norb)
        for Q in range(2 * norb):
            B[Q][iu] = rng.normal(0.0, sigma, len(iu[0]))
            B[Q] = B[Q] + numpy.triu(B[Q], 1).T
        chem = numpy.einsum('Qpq,Qrs->pqrs', B, B)
        phys = chem.transpose(0, 2, 1, 3)          # <pq|rs> = (pr|qs)
        n = 2 * norb
        V = numpy.zeros((n, n, n, n))
        for a in (0, norb):
            for b in (0, norb):
                V[a:a + norb, b:b + norb, a:a + norb, b:b + norb] = phys
        self._eri = V - V.transpose(0, 1, 3, 2)
        self._en = numpy.hstack((e, e))

    def verify(self, T, mu):
        return T == self.T and mu == self.mu

    def has_g(self):
        return True

    def has_u(self):
        return False

    def has_r(self):
        return False

    def const_energy(self):
        return 0.0

    def g_energies_tot(self):
        return self._en.copy()

    def g_fock_tot(self):
        return numpy.diag(self._en)

    def g_aint_tot(self):
        return self._eri

    def get_mp1(self):
        fo = ft_utils.ff(1.0 / self.T, self._en, self.mu)
        return 0.5 * numpy.einsum('ijij,i,j->', self._eri, fo, fo)

    def g_mp1_den(self):
        """beta dE1/dx_p with x_p = beta (e_p - mu) (the content of
        kelvin/ueg_system.py:204-215 for this E1)."""
        beta = 1.0 / self.T
        fo = ft_utils.ff(beta, self._en, self.mu)
        fv = ft_utils.ffv(beta, self._en, self.mu)
        return -beta * numpy.einsum('ijij,i,j->i', self._eri, fo * fv, fo)

    def g_fock_d_den(self):
        """The Fock matrix diag(e) does not depend on the occupations."""
        n = self._en.shape[0]
        return numpy.zeros((n, n, n))


class ArraySystem(System):
    """Provider from explicit arrays: one-body h, antisymmetrised eri,
    reference energies en.  Fock = h + sum_r <pr||qr> fo_r and
    E1 = sum_p (h_pp - e_p) fo_p + 1/2 sum_pq <pq||pq> fo_p fo_q (the
    first-order grand potential of kelvin/scf_system.py:54-68)."""

    def __init__(self, T, mu, h, eri, en, const=0.0):
        self.T, self.mu = T, mu
        self.h = numpy.asarray(h, dtype=float)
        self._eri = numpy.asarray(eri, dtype=float)
        self._en = numpy.asarray(en, dtype=float)
        self._const = const

    def verify(self, T, mu):
        return T == self.T and mu == self.mu

    def has_g(self):
        return True

    def has_u(self):
        return False

    def has_r(self):
        return False

    def const_energy(self):
        return self._const

    def g_energies_tot(self):
        return self._en.copy()

    def _fo(self):
        return ft_utils.ff(1.0 / self.T, self._en, self.mu)

    def g_fock_tot(self):
        return self.h + numpy.einsum('prqr,r->pq', self._eri, self._fo())

    def g_aint_tot(self):
        return self._eri

    def get_mp1(self):
        fo = self._fo()
        return (numpy.dot(numpy.diag(self.h) - self._en, fo)
                + 0.5 * numpy.einsum('ijij,i,j->', self._eri, fo, fo))

    def g_mp1_den(self):
        """beta dE1/dx_p with x_p = beta (e_p - mu)."""
        beta = 1.0 / self.T
        fo = self._fo()
        fv = ft_utils.ffv(beta, self._en, self.mu)
        return -beta * fo * fv * ((numpy.diag(self.h) - self._en)
                                  + numpy.einsum('ijij,j->i', self._eri, fo))

    def g_fock_d_den(self):
        """[p,q,k] = <pk||qk> fo_k fv_k (kelvin/ueg_system.py:498-510 for
        this mean-field Fock matrix)."""
        beta = 1.0 / self.T
        fo = self._fo()
        fv = ft_utils.ffv(beta, self._en, self.mu)
        return numpy.einsum('piqi,i->pqi', self._eri, fo * fv)
