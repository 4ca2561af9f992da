"""Brute-force Fock-space checks (TEST INFRASTRUCTURE ONLY; see oracle/__init__).

Two exact references that need nothing but matrices:

``cc_projection``  <Phi_mu| e^{-T} H e^{T} |Phi_0> for mu = singles and
                   doubles, with H = sum h_pq p+q + 1/4 sum <pq||rs> p+q+sr and
                   T = sum t_ai a+i + 1/4 sum t_abij a+b+ji, built from
                   Jordan-Wigner matrices.  This is the CCSD right-hand side
                   that the Stanton-Gauss factorisation must reproduce term by
                   term (any dropped term, wrong sign or transposed index
                   changes it), for arbitrary (even non-Hermitian) f and
                   antisymmetric <pq||rs>.
``ft_fci_omega``   grand potential -T ln Tr exp(-beta (H - mu N)) over the full
                   Fock space (the content of kelvin/fci.py:88-103).
"""
import numpy
import scipy.linalg


def jw_annihilators(nso):
    """Jordan-Wigner annihilation matrices a_p on 2^nso states.

    Basis state index bit p set <=> mode p occupied; the sign counts the
    occupied modes with smaller index.
    """
    dim = 1 << nso
    ops = []
    for p in range(nso):
        a = numpy.zeros((dim, dim))
        for s in range(dim):
            if s >> p & 1:
                sign = -1.0 if bin(s & ((1 << p) - 1)).count("1") % 2 else 1.0
                a[s ^ (1 << p), s] = sign
        ops.append(a)
    return ops


def hamiltonian(h, eri, ops):
    c = [a.T for a in ops]
    n = len(ops)
    H = numpy.zeros_like(ops[0])
    for p in range(n):
        for q in range(n):
            if h[p, q] != 0.0:
                H += h[p, q] * c[p] @ ops[q]
    for p in range(n):
        for q in range(n):
            cpq = c[p] @ c[q]
            for r in range(n):
                for s in range(n):
                    v = eri[p, q, r, s]
                    if v != 0.0:
                        H += 0.25 * v * cpq @ ops[s] @ ops[r]
    return H


def cc_projection(f, eri, t1, t2, nocc):
    """Return (rhs1[a,i], rhs2[a,b,i,j]) = <Phi| e^{-T} H e^{T} |Phi_0>.

    f is the Fock matrix normal-ordered w.r.t. |Phi_0> (occupied modes
    0..nocc-1); the one-body operator is h = f - sum_i <pi||qi>.
    """
    nso = f.shape[0]
    nvir = nso - nocc
    ops = jw_annihilators(nso)
    c = [a.T for a in ops]
    h = f - numpy.einsum('piqi->pq', eri[:, :nocc, :, :nocc])
    H = hamiltonian(h, eri, ops)
    Tm = numpy.zeros_like(H)
    for a in range(nvir):
        for i in range(nocc):
            Tm += t1[a, i] * c[nocc + a] @ ops[i]
    for a in range(nvir):
        for b in range(nvir):
            for i in range(nocc):
                for j in range(nocc):
                    if t2[a, b, i, j] != 0.0:
                        Tm += 0.25 * t2[a, b, i, j] * (
                            c[nocc + a] @ c[nocc + b] @ ops[j] @ ops[i])
    phi0 = numpy.zeros(H.shape[0])
    phi0[(1 << nocc) - 1] = 1.0      # modes 0..nocc-1 occupied, sign +1
    # |Phi_0> = 0+ 1+ ... (nocc-1)+ |vac> has JW sign +1 with this ordering
    psi = scipy.linalg.expm(Tm) @ phi0
    chi = scipy.linalg.expm(-Tm) @ (H @ psi)
    r1 = numpy.zeros((nvir, nocc))
    r2 = numpy.zeros((nvir, nvir, nocc, nocc))
    for a in range(nvir):
        for i in range(nocc):
            r1[a, i] = (c[nocc + a] @ ops[i] @ phi0) @ chi
    for a in range(nvir):
        for b in range(nvir):
            for i in range(nocc):
                for j in range(nocc):
                    st = c[nocc + a] @ c[nocc + b] @ ops[j] @ ops[i] @ phi0
                    r2[a, b, i, j] = st @ chi
    return r1, r2


def ft_fci_omega(h, eri, beta, mu):
    """-1/beta ln Tr exp(-beta (H - mu N)) over all particle numbers."""
    nso = h.shape[0]
    ops = jw_annihilators(nso)
    H = hamiltonian(h, eri, ops)
    Nop = sum(a.T @ a for a in ops)
    w = numpy.linalg.eigvalsh(0.5 * (H + H.T) - mu * Nop)
    wmin = w.min()
    return wmin - numpy.log(numpy.exp(-beta * (w - wmin)).sum()) / beta


def random_antisym(n, rng, hermitian=True):
    """Random real <pq||rs>, antisymmetric in (pq) and (rs); with
    hermitian=True also <pq||rs> = <rs||pq>."""
    x = rng.normal(size=(n, n, n, n))
    x = x - x.transpose(1, 0, 2, 3)
    x = x - x.transpose(0, 1, 3, 2)
    if hermitian:
        x = x + x.transpose(2, 3, 0, 1)
    return x
