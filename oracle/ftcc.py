"""numpy FP64 restatement of kelvin's finite-temperature CCSD hot path.

TEST INFRASTRUCTURE ONLY: this is the oracle.  The product package
``kelvin_amd`` never imports it; tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker.

Every function cites the reference file:line it restates (paths relative to
the reference checkout).  The CCSD residual itself lives in the third-party
package cqcpy (``cc_equations._Stanton``, unversioned, HEAD of
github.com/awhite862/cqcpy per ``.github/workflows/clone_deps.sh:2``), which is
absent here; ``stanton_rhs`` restates the published Stanton-Gauss equations
(J. F. Stanton, J. Gauss, J. D. Watts, R. J. Bartlett, J. Chem. Phys. 94, 4334
(1991); spin-orbital form of Crawford & Schaefer, Rev. Comp. Chem. 14, 33
(2000)) in kelvin's block conventions.

How each piece is pinned (tests/test_oracle_*.py):
  * quadrature rules: exactness of Simpson on cubics, closed forms, the
    finite-difference identity for the ``d_*`` beta-derivative rules.
  * Fermi factors / GP0 / dGP0: closed forms and E0 = beta*sum(dGP0)+mu*N0+G0.
  * UEG integrals: antisymmetry, pair-exchange symmetry, momentum conservation,
    and the end-to-end goldens below.
  * integrators: the constant-source closed form T(tau) = Tbar (1-e^{-tau D})/D
    (up to Simpson error) and agreement of the literal O(ng^2) and scan forms.
  * ``stanton_rhs``: brute-force Fock-space projection
    <Phi_mu| e^{-T} H e^{T} |Phi_0> on random non-Hermitian Hamiltonians
    (oracle/fock_space.py), then end to end.
  * ``ccsd_1rdm``: every element against the derivative of
    ``ft_lagrangian`` with respect to its Fock element, evaluated exactly by
    unit perturbations (the Lagrangian is linear in the Fock blocks) -- the
    identity kelvin/tests/test_ft_ccsd_rdm.py:12-170 checks as a trace;
    ``normal_1rdm`` (less its occupations) against the derivative of the
    Lagrangian with respect to the orbital Fock matrix through
    ``ft_integrals``' square-root weights, also exact by unit perturbations.
  * adjoint (``stanton_rhs_transpose``, ``int_adjoint``, ``int_L1/2``,
    ``ft_lambda_iter``, ``response_blocks``, ``grid_derivative``,
    ``ft_ccsd_esn``): the dot-product identity <Lam, J X> = <J^T Lam, X>
    (exact: the residual is quadratic), the reverse integration as the
    transpose of int_tbar1/2, the stationarity of the Lagrangian at the
    converged lambdas (kelvin/tests/test_ft_lambda.py:16-133), kelvin's
    E, S, N of the G4 example, and central differences of full solves.
  * ``ft_active_integrals``: every block equals ``ft_integrals``' block
    restricted to the kept lists; keeping every orbital reproduces the
    full-space solve.
  * systems: ``UEGSystem`` by its symmetries and the goldens;
    ``PUEGSystem`` by the reference's polarised-gas grand potential
    (kelvin/tests/test_ft_ccsd.py:24, whose printed value is this oracle's
    iteration-10 energy to 1e-12); ``DFModelSystem`` by the positive
    semidefinite, rank-limited (pr|qs) block and the spin zero pattern.
  * TD-CCSD (oracle/td_ccsd.py): the order of each propagator on an exactly
    solvable problem, and agreement with the fixed point.
  * end to end: the reference's golden UEG grand potentials
    (kelvin/tests/test_ft_ccsd.py:23 with :119-133, Omega_cc=-0.010750238811
    at 1e-9) and the per-iteration energies printed in
    examples/ueg_ft_cc_compare.out:10-26 and examples/ueg_ft_ccsd_ESN.out:9-25.
"""
import logging
import numpy

einsum = numpy.einsum


def _es(spec, *ops):
    return numpy.einsum(spec, *ops, optimize=True)


# ---------------------------------------------------------------------------
# Fermi-Dirac factors and free-fermion grand potential (cqcpy ft_utils; call
# sites kelvin/ccsd.py:627-629,637-638, kelvin/cc_utils.py:571-574).  Closed
# forms, written so that tiny occupations keep their relative precision.
# ---------------------------------------------------------------------------
def ff(beta, en, mu):
    """Occupied Fermi factor 1/(1+exp(beta(e-mu)))."""
    x = beta * (numpy.asarray(en, dtype=float) - mu)
    out = numpy.empty_like(x)
    p = x > 0
    ex = numpy.exp(-x[p])
    out[p] = ex / (1.0 + ex)
    out[~p] = 1.0 / (1.0 + numpy.exp(x[~p]))
    return out


def ffv(beta, en, mu):
    """Virtual Fermi factor 1 - ff = 1/(1+exp(-beta(e-mu)))."""
    return ff(beta, -(numpy.asarray(en, dtype=float) - mu), 0.0)


def GP0(beta, en, mu):
    """Per-orbital free-fermion grand potential -(1/beta) ln(1+e^{-beta(e-mu)})."""
    x = beta * (numpy.asarray(en, dtype=float) - mu)
    # ln(1+e^{-x}) = max(-x,0) + log1p(e^{-|x|})
    return -(numpy.maximum(-x, 0.0) + numpy.log1p(numpy.exp(-numpy.abs(x)))) / beta


def dGP0(beta, en, mu):
    """d GP0 / d beta per orbital (used by kelvin/ccsd.py:177-180)."""
    e = numpy.asarray(en, dtype=float)
    return ((e - mu) * ff(beta, e, mu) - GP0(beta, e, mu)) / beta


def D1(ev, eo):
    """D1[a,i] = e_a - e_i (cqcpy utils.D1, call site kelvin/ccsd.py:633)."""
    return ev[:, None] - eo[None, :]


def D2(ev, eo):
    """D2[a,b,i,j] = e_a + e_b - e_i - e_j (cqcpy utils.D2, kelvin/ccsd.py:634)."""
    return (ev[:, None, None, None] + ev[None, :, None, None]
            - eo[None, None, :, None] - eo[None, None, None, :])


# ---------------------------------------------------------------------------
# Quadrature rules: kelvin/quadrature.py:17-256
# ---------------------------------------------------------------------------
def get_G(ng, delta):
    """kelvin/quadrature.py:32-41 (Simpson weights for int_0^{t_y})."""
    G = numpy.zeros((ng, ng))
    if ng > 1:
        G[1, 0] = G[1, 1] = 0.5 * delta
    for y in range(2, ng):
        G[y] = G[y - 2]
        G[y, y - 2] += delta / 3.0
        G[y, y - 1] += 4.0 * delta / 3.0
        G[y, y] += delta / 3.0
    return G


def get_gint(ng, delta):
    """kelvin/quadrature.py:44-60 (Simpson weights for int_0^beta)."""
    g = numpy.zeros(ng)
    if ng % 2 == 0:
        g[0] += 0.5 * delta
        g[1] += 0.5 * delta
        for y in range(3, ng, 2):
            g[y - 2] += delta / 3.0
            g[y - 1] += 4.0 * delta / 3.0
            g[y] += delta / 3.0
    else:
        for y in range(2, ng, 2):
            g[y - 2] += delta / 3.0
            g[y - 1] += 4.0 * delta / 3.0
            g[y] += delta / 3.0
    return g


def get_G_midpoint(ng, delta):
    """kelvin/quadrature.py:17-23."""
    G = numpy.zeros((ng, ng))
    G[0, 0] = delta
    for y in range(1, ng):
        G[y] = G[y - 1]
        G[y, y] += delta
    return G


def get_gL(ng, delta):
    """kelvin/quadrature.py:63-67."""
    g = numpy.zeros(ng)
    g[:ng - 1] = delta
    return g


def get_GL(ng, delta):
    """kelvin/quadrature.py:70-75."""
    G = numpy.zeros((ng, ng))
    for i in range(ng):
        for j in range(i - 1):
            G[i, j] = delta
    return G


def _mapped(ng, beta, delta, si, jac):
    g = beta * get_gint(ng, delta) * jac
    G = get_G(ng, delta)
    for i in range(ng):
        G[i] = beta * G[i] * jac
    return g, G


def ft_quad(ng, beta, quad):
    """kelvin/quadrature.py:215-235 with the rules at :78-212."""
    if quad == 'lin':                                   # :102-107
        delta = beta / (ng - 1.0)
        ti = numpy.arange(ng) * delta
        return ti, get_gint(ng, delta), get_G(ng, delta)
    if quad == 'ln':                                    # :119-129
        e = numpy.exp(1)
        delta = (e - 1.0) / (ng - 1.0)
        si = numpy.arange(ng) * delta + 1.0
        g = beta * get_gint(ng, delta) / si
        G = get_G(ng, delta)
        for i in range(ng):
            G[i] = beta * G[i] / si
        return beta * numpy.log(si), g, G
    if quad == 'sin':                                   # :144-153
        delta = numpy.pi / (ng - 1.0)
        si = numpy.arange(ng) * delta - numpy.pi / 2
        g, G = _mapped(ng, beta, delta, si, numpy.cos(si) / 2.0)
        return beta * (numpy.sin(si) + 1.0) / 2.0, g, G
    if quad == 'exp':                                   # :166-176
        delta = numpy.log(2.0) / (ng - 1.0)
        si = numpy.arange(ng) * delta
        g, G = _mapped(ng, beta, delta, si, numpy.exp(si))
        return beta * (numpy.exp(si) - 1.0), g, G
    if quad in ('quad', 'cub', 'quar'):                 # :190-199
        p = {'quad': 2, 'cub': 3, 'quar': 4}[quad]
        delta = 1.0 / (ng - 1.0)
        si = numpy.arange(ng) * delta
        jac = (numpy.power(si, p - 1) * float(p) + 1.0) / 2.0
        g, G = _mapped(ng, beta, delta, si, jac)
        return beta * (numpy.power(si, p) + si) / 2.0, g, G
    if quad == 'mid':                                   # :94-99
        delta = beta / ng
        ti = numpy.arange(ng) * delta + delta / 2
        return ti, numpy.full(ng, delta), get_G_midpoint(ng, delta)
    if quad == 'L':                                     # :78-83
        delta = beta / (ng - 1.0)
        ti = numpy.arange(ng) * delta
        return ti, get_gL(ng, delta), get_GL(ng, delta)
    raise Exception("Unrecognized quadrature rule: {}".format(quad))


def d_ft_quad(ng, beta, quad):
    """kelvin/quadrature.py:238-256: d/dbeta of (g, G) at fixed mapped grid."""
    if quad == 'lin':                                   # :110-116
        delta = beta / (ng - 1.0)
        return get_gint(ng, delta / beta), get_G(ng, delta / beta)
    if quad == 'L':                                     # :86-91
        delta = beta / (ng - 1.0)
        return get_gL(ng, delta / beta), get_GL(ng, delta / beta)
    if quad in ('ln', 'sin', 'exp', 'quad', 'cub', 'quar'):
        # the mapped rules are linear in beta: d/dbeta = value at beta=1
        ti, g, G = ft_quad(ng, 1.0, quad)
        return g, G
    raise Exception("Unrecognized quadrature rule: {}".format(quad))


# ---------------------------------------------------------------------------
# exp(-tau D) integrators: kelvin/quadrature.py:292-317 (literal O(ng^2) form)
# ---------------------------------------------------------------------------
def int_tbar1(ng, t1bar, ti, D1, G):
    """kelvin/quadrature.py:292-303."""
    t1_out = numpy.zeros(t1bar.shape, dtype=t1bar.dtype)
    dt = numpy.zeros(ng)
    for y in range(ng):
        for i in range(y):
            dt[i] = ti[i] - ti[y]
        gtemp = numpy.exp(dt[:, None, None] * D1[None, :, :])
        t1_out[y] = einsum('x,xai->ai', G[y], gtemp * t1bar)
    return t1_out


def int_tbar2(ng, t2bar, ti, D2, G):
    """kelvin/quadrature.py:306-317."""
    t2_out = numpy.zeros(t2bar.shape, dtype=t2bar.dtype)
    dt = numpy.zeros(ng)
    for y in range(ng):
        for i in range(y):
            dt[i] = ti[i] - ti[y]
        gtemp = numpy.exp(dt[:, None, None, None, None] * D2[None])
        t2_out[y] = einsum('x,xabij->abij', G[y], gtemp * t2bar)
    return t2_out


def int_tbar_scan(ng, tbar, ti, D, G):
    """Same result as int_tbar1/2, one slice at a time (memory O(ng*size)).

    Mirrors the reference's dt convention exactly: row y only ever sets
    dt[x] for x < y, so for x >= y the propagator is exp(0) = 1 and any
    weight G[y,x] on or above the diagonal multiplies Tbar[x] unpropagated
    (kelvin/quadrature.py:296-299).
    """
    out = numpy.zeros_like(tbar)
    dt = numpy.zeros(ng)
    for y in range(ng):
        for i in range(y):
            dt[i] = ti[i] - ti[y]
        acc = numpy.zeros_like(tbar[0])
        for x in range(ng):
            if G[y, x] != 0.0:
                acc += G[y, x] * numpy.exp(dt[x] * D) * tbar[x]
        out[y] = acc
    return out


# ---------------------------------------------------------------------------
# Fermi-weighted integrals: kelvin/cc_utils.py:569-602 and :825-858
# ---------------------------------------------------------------------------
class Blocks(dict):
    """Attribute bag standing in for cqcpy one_e_blocks / two_e_blocks."""
    __getattr__ = dict.__getitem__


def _weigh(eri, idx, w):
    (p, q, r, s), (wp, wq, wr, ws) = idx, w
    return (eri[numpy.ix_(p, q, r, s)] * wp[:, None, None, None]
            * wq[None, :, None, None] * wr[None, None, :, None]
            * ws[None, None, None, :])


def ft_integrals(fock, eri, en, beta, mu):
    """kelvin/cc_utils.py:569-602 (fock = sys.g_fock_tot(), eri = g_aint_tot())."""
    n = en.shape[0]
    iall = numpy.arange(n)
    return ft_active_integrals(fock, eri, en, ff(beta, en, mu),
                               ffv(beta, en, mu), iall, iall)


def ft_active_integrals(fock, eri, en, focc, fvir, iocc, ivir):
    """kelvin/cc_utils.py:825-858 (focc/fvir already restricted to iocc/ivir)."""
    fmo = fock - numpy.diag(en)                                  # :577-578
    so = numpy.sqrt(numpy.asarray(focc, dtype=float))
    sv = numpy.sqrt(numpy.asarray(fvir, dtype=float))
    o, v = numpy.asarray(iocc), numpy.asarray(ivir)
    F = Blocks(
        oo=fmo[numpy.ix_(o, o)] * so[:, None] * so[None, :],
        ov=fmo[numpy.ix_(o, v)] * so[:, None] * sv[None, :],
        vo=fmo[numpy.ix_(v, o)] * sv[:, None] * so[None, :],
        vv=fmo[numpy.ix_(v, v)] * sv[:, None] * sv[None, :])
    W = {'o': (o, so), 'v': (v, sv)}
    I = Blocks()
    for name in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo',
                 'oovv', 'vooo', 'ooov', 'oooo'):
        idx = tuple(W[c][0] for c in name)
        w = tuple(W[c][1] for c in name)
        I[name] = _weigh(eri, idx, w)
    return F, I


# ---------------------------------------------------------------------------
# Stanton-Gauss CCSD right-hand side (cqcpy cc_equations._Stanton, called at
# kelvin/ft_cc_equations.py:105-107 with fac=-1).  All indices run over the
# occupied (m,n,i,j) or virtual (a,b,e,f) lists, which at finite T are both
# the full orbital set.  Layouts: t1[a,i], t2[a,b,i,j].
#
# Integral map (antisymmetry <pq||rs> = -<qp||rs> = -<pq||sr> = <qp||sr>):
#   <ma||fe> = -vovv[a,m,f,e]   <mb||ej> = -vovo[b,m,e,j]   <na||if> = vovo[a,n,f,i]
#   <mn||ej> = -ooov[m,n,j,e]   <nm||ei> = -ooov[n,m,i,e]   <mb||ij> = -vooo[b,m,i,j]
# ---------------------------------------------------------------------------
def stanton_rhs(F, I, t1, t2):
    """Return (rhs1, rhs2) = (f_ai + R1(T), <ab||ij> + R2(T)).

    kelvin's residual is Tbar1 = -rhs1, Tbar2 = -rhs2
    (kelvin/ft_cc_equations.py:101-107: T1new = -F.vo, then _Stanton with
    fac=-1 adds -R1).  The Fock blocks are used in full (no (1-delta) mask):
    the caller has already removed diag(en) (kelvin/cc_utils.py:577-578).
    """
    es = _es
    vovv, ooov, oovv = I.vovv, I.ooov, I.oovv
    tt = es('ai,bj->abij', t1, t1)
    tt = tt - tt.transpose(1, 0, 2, 3)
    tau = t2 + tt
    taut = t2 + 0.5 * tt

    # one-particle intermediates
    Fme = F.ov + es('fn,mnef->me', t1, oovv)
    Fae = (F.vv - 0.5 * es('me,am->ae', F.ov, t1)
           - es('fm,amfe->ae', t1, vovv)
           - 0.5 * es('afmn,mnef->ae', taut, oovv))
    Fmi = (F.oo + 0.5 * es('ei,me->mi', t1, F.ov)
           + es('en,mnie->mi', t1, ooov)
           + 0.5 * es('efin,mnef->mi', taut, oovv))

    # two-particle intermediates
    tmp = es('ej,mnie->mnij', t1, ooov)
    Wmnij = (I.oooo + tmp - tmp.transpose(0, 1, 3, 2)
             + 0.25 * es('efij,mnef->mnij', tau, oovv))
    tmp = es('bm,amef->abef', t1, vovv)
    Wabef = (I.vvvv - tmp + tmp.transpose(1, 0, 2, 3)
             + 0.25 * es('abmn,mnef->abef', tau, oovv))
    ovvo = -I.vovo.transpose(1, 0, 2, 3)             # <mb||ej> as [m,b,e,j]
    Wmbej = (ovvo - es('fj,bmef->mbej', t1, vovv)
             + es('bn,mnje->mbej', t1, ooov)
             - 0.5 * es('fbjn,mnef->mbej', t2, oovv)
             - es('fj,bn,mnef->mbej', t1, t1, oovv))

    # singles
    r1 = (F.vo + es('ei,ae->ai', t1, Fae) - es('am,mi->ai', t1, Fmi)
          + es('aeim,me->ai', t2, Fme)
          - es('fn,anfi->ai', t1, I.vovo)
          + 0.5 * es('efim,amef->ai', t2, vovv)
          + 0.5 * es('aemn,nmie->ai', t2, ooov))

    # doubles
    r2 = I.vvoo.copy()
    Fbe = Fae - 0.5 * es('bm,me->be', t1, Fme)
    tmp = es('aeij,be->abij', t2, Fbe)
    r2 += tmp - tmp.transpose(1, 0, 2, 3)
    Fmj = Fmi + 0.5 * es('ej,me->mj', t1, Fme)
    tmp = es('abim,mj->abij', t2, Fmj)
    r2 -= tmp - tmp.transpose(0, 1, 3, 2)
    r2 += 0.5 * es('abmn,mnij->abij', tau, Wmnij)
    r2 += 0.5 * es('efij,abef->abij', tau, Wabef)
    tmp = (es('aeim,mbej->abij', t2, Wmbej)
           - es('ei,am,mbej->abij', t1, t1, ovvo))
    r2 += (tmp - tmp.transpose(1, 0, 2, 3) - tmp.transpose(0, 1, 3, 2)
           + tmp.transpose(1, 0, 3, 2))
    tmp = es('ei,abej->abij', t1, I.vvvo)
    r2 += tmp - tmp.transpose(0, 1, 3, 2)
    tmp = es('am,bmij->abij', t1, I.vooo)
    r2 += tmp - tmp.transpose(1, 0, 2, 3)
    return r1, r2


def ccsd_stanton(F, I, T1old, T2old, D1, D2, ti, ng, G):
    """kelvin/ft_cc_equations.py:96-113."""
    T1new = numpy.empty_like(T1old)
    T2new = numpy.empty_like(T2old)
    for y in range(ng):
        r1, r2 = stanton_rhs(F, I, T1old[y], T2old[y])
        T1new[y] = -r1
        T2new[y] = -r2
    return int_tbar1(ng, T1new, ti, D1, G), int_tbar2(ng, T2new, ti, D2, G)


def residual_tbar(F, I, T1, T2):
    """Pre-integration residual at every slice: (Tbar1, Tbar2)."""
    ng = T1.shape[0]
    T1b = numpy.empty_like(T1)
    T2b = numpy.empty_like(T2)
    for y in range(ng):
        r1, r2 = stanton_rhs(F, I, T1[y], T2[y])
        T1b[y], T2b[y] = -r1, -r2
    return T1b, T2b


# ---------------------------------------------------------------------------
# Adjoint of the residual (the Lambda equations).  kelvin solves them with
# cqcpy cc_equations._Lambda_opt (absent, called at
# kelvin/ft_cc_equations.py:385-409 with fac=-1), whose content is the
# transpose of the Jacobian of the Stanton-Gauss residual contracted with the
# integrated lambda amplitudes.  stanton_rhs_transpose restates that operator
# as the exact reverse-mode derivative of stanton_rhs, term by term.
# ---------------------------------------------------------------------------
def stanton_rhs_transpose(F, I, t1, t2, v1, v2):
    """Return (g1, g2) with <g, w> = d/dh <v, stanton_rhs(t + h w)> at h = 0
    for every direction w (all n^4 components of t2 independent: no
    antisymmetry is assumed of v2 or w2).  Layouts as stanton_rhs."""
    es = _es
    vovv, ooov, oovv, vovo = I.vovv, I.ooov, I.oovv, I.vovo
    fov = F.ov
    # forward intermediates, as stanton_rhs
    tt0 = es('ai,bj->abij', t1, t1)
    tt = tt0 - tt0.transpose(1, 0, 2, 3)
    tau = t2 + tt
    taut = t2 + 0.5 * tt
    Fme = fov + es('fn,mnef->me', t1, oovv)
    Fae = (F.vv - 0.5 * es('me,am->ae', fov, t1)
           - es('fm,amfe->ae', t1, vovv)
           - 0.5 * es('afmn,mnef->ae', taut, oovv))
    Fmi = (F.oo + 0.5 * es('ei,me->mi', t1, fov)
           + es('en,mnie->mi', t1, ooov)
           + 0.5 * es('efin,mnef->mi', taut, oovv))
    tmp = es('ej,mnie->mnij', t1, ooov)
    Wmnij = (I.oooo + tmp - tmp.transpose(0, 1, 3, 2)
             + 0.25 * es('efij,mnef->mnij', tau, oovv))
    tmp = es('bm,amef->abef', t1, vovv)
    Wabef = (I.vvvv - tmp + tmp.transpose(1, 0, 2, 3)
             + 0.25 * es('abmn,mnef->abef', tau, oovv))
    ovvo = -vovo.transpose(1, 0, 2, 3)
    Wmbej = (ovvo - es('fj,bmef->mbej', t1, vovv)
             + es('bn,mnje->mbej', t1, ooov)
             - 0.5 * es('fbjn,mnef->mbej', t2, oovv)
             - es('fj,bn,mnef->mbej', t1, t1, oovv))
    Fbe = Fae - 0.5 * es('bm,me->be', t1, Fme)
    Fmj = Fmi + 0.5 * es('ej,me->mj', t1, Fme)

    def pab(x):
        return x.transpose(1, 0, 2, 3)

    def pij(x):
        return x.transpose(0, 1, 3, 2)

    g1 = numpy.zeros_like(t1)
    g2 = numpy.zeros_like(t2)
    # doubles terms: transposes of the antisymmetrisers
    gX3 = v2 - pab(v2)
    gX4 = -(v2 - pij(v2))
    gX5 = v2 - pab(v2) - pij(v2) + pab(pij(v2))
    gX6 = v2 - pij(v2)
    gX7 = v2 - pab(v2)
    g1 += es('abij,bmij->am', gX7, I.vooo)          # t1[a,m] vooo[b,m,i,j]
    g1 += es('abij,abej->ei', gX6, I.vvvo)          # t1[e,i] vvvo[a,b,e,j]
    g2 += es('abij,mbej->aeim', gX5, Wmbej)         # t2[a,e,i,m] Wmbej
    gW = es('aeim,abij->mbej', t2, gX5)
    Z = es('abij,mbej->aime', gX5, ovvo)            # -t1[e,i] t1[a,m] ovvo
    g1 -= es('aime,am->ei', Z, t1)
    g1 -= es('aime,ei->am', Z, t1)
    gtau = 0.5 * es('abij,abef->efij', v2, Wabef)   # 1/2 tau.Wabef
    gWabef = 0.5 * es('abij,efij->abef', v2, tau)
    gtau += 0.5 * es('abij,mnij->abmn', v2, Wmnij)  # 1/2 tau.Wmnij
    gWmnij = 0.5 * es('abmn,abij->mnij', tau, v2)
    g2 += es('abij,mj->abim', gX4, Fmj)             # t2[a,b,i,m] Fmj
    gFmj = es('abim,abij->mj', t2, gX4)
    g2 += es('abij,be->aeij', gX3, Fbe)             # t2[a,e,i,j] Fbe
    gFbe = es('aeij,abij->be', t2, gX3)
    # Fmj = Fmi + 1/2 t1 Fme ; Fbe = Fae - 1/2 t1 Fme
    gFmi = gFmj.copy()
    g1 += 0.5 * es('mj,me->ej', gFmj, Fme)
    gFme = 0.5 * es('mj,ej->me', gFmj, t1)
    gFae = gFbe.copy()
    g1 -= 0.5 * es('be,me->bm', gFbe, Fme)
    gFme -= 0.5 * es('be,bm->me', gFbe, t1)
    # singles terms
    gFae += es('ai,ei->ae', v1, t1)
    g1 += es('ai,ae->ei', v1, Fae)
    gFmi -= es('ai,am->mi', v1, t1)
    g1 -= es('ai,mi->am', v1, Fmi)
    g2 += es('ai,me->aeim', v1, Fme)
    gFme += es('ai,aeim->me', v1, t2)
    g1 -= es('ai,anfi->fn', v1, vovo)
    g2 += 0.5 * es('ai,amef->efim', v1, vovv)
    g2 += 0.5 * es('ai,nmie->aemn', v1, ooov)
    # Wmbej
    g1 -= es('mbej,bmef->fj', gW, vovv)
    g1 += es('mbej,mnje->bn', gW, ooov)
    g2 -= 0.5 * es('mbej,mnef->fbjn', gW, oovv)
    Q = es('mbej,mnef->bjnf', gW, oovv)
    g1 -= es('bjnf,bn->fj', Q, t1)
    g1 -= es('bjnf,fj->bn', Q, t1)
    # Wabef: -t1[b,m] vovv[a,m,e,f] + t1[a,m] vovv[b,m,e,f] + 1/4 tau.oovv
    gX2 = -gWabef + pab(gWabef)
    g1 += es('abef,amef->bm', gX2, vovv)
    gtau += 0.25 * es('abef,mnef->abmn', gWabef, oovv)
    # Wmnij: t1[e,j] ooov[m,n,i,e] - (i <-> j) + 1/4 tau.oovv
    gX1 = gWmnij - pij(gWmnij)
    g1 += es('mnij,mnie->ej', gX1, ooov)
    gtau += 0.25 * es('mnij,mnef->efij', gWmnij, oovv)
    # Fmi
    g1 += 0.5 * es('mi,me->ei', gFmi, fov)
    g1 += es('mi,mnie->en', gFmi, ooov)
    gtaut = 0.5 * es('mi,mnef->efin', gFmi, oovv)
    # Fae
    g1 -= 0.5 * es('ae,me->am', gFae, fov)
    g1 -= es('ae,amfe->fm', gFae, vovv)
    gtaut -= 0.5 * es('ae,mnef->afmn', gFae, oovv)
    # Fme
    g1 += es('me,mnef->fn', gFme, oovv)
    # tau = t2 + tt, taut = t2 + tt/2, tt = t1 t1 - (a <-> b)
    g2 += gtau + gtaut
    gtt = gtau + 0.5 * gtaut
    gtt0 = gtt - pab(gtt)
    g1 += es('abij,bj->ai', gtt0, t1) + es('abij,ai->bj', gtt0, t1)
    return g1, g2


def antisymmetrize(x2):
    """Orthogonal projection of (..., a, b, i, j) onto tensors antisymmetric
    in (a, b) and in (i, j)."""
    y = x2 - numpy.swapaxes(x2, -4, -3)
    y = y - numpy.swapaxes(y, -2, -1)
    return 0.25 * y


def int_adjoint(ng, lam, ti, D, G):
    """Transpose of int_tbar1/int_tbar2 (kelvin/quadrature.py:292-317) in
    the slice index: out[x] = sum_y G[y,x] P_yx lam[y], with the integrator's
    propagator P_yx = exp((ti[x] - ti[y]) D) for x < y and 1 for x >= y."""
    out = numpy.zeros_like(lam)
    for x in range(ng):
        acc = numpy.zeros_like(lam[0])
        for y in range(ng):
            if G[y, x] == 0.0:
                continue
            w = numpy.exp((ti[x] - ti[y]) * D) if x < y else 1.0
            acc += G[y, x] * w * lam[y]
        out[x] = acc
    return out


def int_L1(ng, L1old, ti, D1, g, G):
    """kelvin/quadrature.py:320-332 (L1 layout [y,i,a])."""
    L1_out = numpy.zeros(L1old.shape, dtype=L1old.dtype)
    for s in range(ng):
        dt = numpy.zeros(ng)
        for y in range(s, ng):
            dt[y] = ti[s] - ti[y]
        gtemp = numpy.exp(dt[:, None, None] * D1[None, :, :])
        L1temp = einsum('yai,yia->yia', gtemp, L1old)
        L1_out[s] = einsum('y,yia,y->ia', g, L1temp, G[:, s]) / g[s]
    return L1_out


def int_L2(ng, L2old, ti, D2, g, G):
    """kelvin/quadrature.py:335-345 (L2 layout [y,i,j,a,b])."""
    L2_out = numpy.zeros(L2old.shape, dtype=L2old.dtype)
    for s in range(ng):
        dt = numpy.zeros(ng)
        for y in range(s, ng):
            dt[y] = ti[s] - ti[y]
        gtemp = numpy.exp(dt[:, None, None, None, None] * D2[None])
        L2temp = einsum('yabij,yijab->yijab', gtemp, L2old)
        L2_out[s] = einsum('y,yijab,y->ijab', g, L2temp, G[:, s]) / g[s]
    return L2_out


def ccsd_lambda_guess(F, I, T1, beta, ng):
    """kelvin/ft_cc_equations.py:502-512 with cqcpy _LS_TS (the T1 term of
    the energy's derivative, +fac * sum_jb oovv[i,j,a,b] t1[b,j])."""
    L1 = numpy.broadcast_to(F.ov / beta, (ng,) + F.ov.shape).copy()
    L2 = numpy.broadcast_to(I.oovv / beta, (ng,) + I.oovv.shape).copy()
    for y in range(ng):
        L1[y] += einsum('ijab,bj->ia', I.oovv, T1[y]) / beta
    return L1, L2


def ccsd_lambda_opt(F, I, T1, T2, L1old, L2old, D1, D2, ti, ng, g, G, beta):
    """kelvin/ft_cc_equations.py:385-409 in kelvin's lambda convention
    (L1[y,i,a], L2[y,i,j,a,b]; the derivative of Omega_cc along a change dT
    is -(1/beta) sum_y g_y [L1.dT1 + 1/4 L2.dT2], kelvin/ccsd.py:1113-1146):
    L = -dE_y/dT_y - J_y^T (int_L(L)) with the Jacobian transpose of
    stanton_rhs_transpose, the doubles projected onto antisymmetric
    tensors."""
    L1int = int_L1(ng, L1old, ti, D1, g, G)
    L2int = int_L2(ng, L2old, ti, D2, g, G)
    L1 = numpy.empty_like(L1old)
    L2 = numpy.empty_like(L2old)
    for y in range(ng):
        v1 = L1int[y].T
        v2 = 0.25 * L2int[y].transpose(2, 3, 0, 1)
        g1, g2 = stanton_rhs_transpose(F, I, T1[y], T2[y], v1, v2)
        g2 = antisymmetrize(g2)
        L1[y] = (-F.ov - einsum('ijab,bj->ia', I.oovv, T1[y]) - g1.T)
        L2[y] = -I.oovv - 4.0 * g2.transpose(2, 3, 0, 1)
    return L1, L2


def ft_lambda_iter(L1old, L2old, T1, T2, F, I, D1, D2, g, G, beta, ng, ti,
                   conv, history=None):
    """kelvin/cc_utils.py:414-480: damped iteration of ccsd_lambda_opt with
    norms ||L|| + 0.1 and the test res1 + res2 < tconv."""
    thresh, max_iter, alpha = conv["tconv"], conv["max_iter"], conv["damp"]
    i, converged = 0, False
    nl1 = numpy.linalg.norm(L1old) + 0.1
    nl2 = numpy.linalg.norm(L2old) + 0.1
    while i < max_iter and not converged:
        L1, L2 = ccsd_lambda_opt(F, I, T1, T2, L1old, L2old, D1, D2, ti, ng,
                                 g, G, beta)
        res1 = numpy.linalg.norm(L1 - L1old) / nl1
        res2 = numpy.linalg.norm(L2 - L2old) / nl2
        L1old = alpha * L1old + (1.0 - alpha) * L1
        L2old = alpha * L2old + (1.0 - alpha) * L2
        nl1 = numpy.linalg.norm(L1old) + 0.1
        nl2 = numpy.linalg.norm(L2old) + 0.1
        if history is not None:
            history.append(res1 + res2)
        i += 1
        converged = res1 + res2 < thresh
    return L1old, L2old, converged, i


def response_blocks(F, I, dfock, fo, fv, dx, iocc, ivir):
    """Derivative of the Fermi-weighted blocks F, I (kelvin/cc_utils.py:
    569-602, :825-858, over the orbital lists iocc / ivir) when
    x_p = beta (e_p - mu) moves by dx_p (all n orbitals) and the bare Fock
    matrix by dfock (n x n).  With d ln sqrt(fo_p) = -fv_p dx_p / 2 and
    d ln sqrt(fv_p) = fo_p dx_p / 2 a block w x y z changes by
    I * (c^w_p + c^x_q + c^y_r + c^z_s).  These are the quantities kelvin
    contracts with its response densities in g_d_on_oo, g_d_on_vv (the
    dso = fv, dsv = fo factors) and g_Fd_on (kelvin/cc_utils.py:1628-1700)."""
    o, v = numpy.asarray(iocc), numpy.asarray(ivir)
    so, sv = numpy.sqrt(fo[o]), numpy.sqrt(fv[v])
    c = {'o': -0.5 * fv[o] * dx[o], 'v': 0.5 * fo[v] * dx[v]}
    s_ = {'o': so, 'v': sv}
    idx = {'o': o, 'v': v}
    dF = Blocks()
    for name in ('oo', 'ov', 'vo', 'vv'):
        x, y = name
        dF[name] = (dfock[numpy.ix_(idx[x], idx[y])] * s_[x][:, None]
                    * s_[y][None, :]
                    + F[name] * (c[x][:, None] + c[y][None, :]))
    dI = Blocks()
    for name in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo', 'oovv', 'vooo',
                 'ooov', 'oooo'):
        w = (c[name[0]][:, None, None, None] + c[name[1]][None, :, None, None]
             + c[name[2]][None, None, :, None]
             + c[name[3]][None, None, None, :])
        dI[name] = I[name] * w
    return dF, dI


def grid_derivative(Lam1, Lam2, Tb1, Tb2, ti, G, Gd, D1, D2, beta):
    """sum over y, x of Lam_y . (Gd[y,x] P_yx - G[y,x] (ti_y - ti_x)/beta D
    P_yx) Tbar_x: the change of the integrated amplitudes with beta through
    the weights and the propagators (ti scales with beta), contracted with
    the adjoint amplitudes (kelvin/ccsd.py:1113-1146 in the unscaled
    convention; P as in int_adjoint)."""
    ng = ti.shape[0]
    out = 0.0
    for Lam, Tb, D in ((Lam1, Tb1, D1), (Lam2, Tb2, D2)):
        for y in range(ng):
            for x in range(ng):
                if G[y, x] == 0.0 and Gd[y, x] == 0.0:
                    continue
                P = numpy.exp((ti[x] - ti[y]) * D) if x < y else 1.0
                K = (Gd[y, x] - G[y, x] * (ti[y] - ti[x]) / beta * D) * P
                out += numpy.sum(Lam[y] * K * Tb[x])
    return out


def ft_ccsd_esn(sys, T, mu, ngrid=10, damp=0.0, max_iter=40, econv=1e-8,
                tconv=None, quad='lin', athresh=0.0, gderiv=True):
    """E, S, N of the FT-CCSD grand potential (kelvin/ccsd.py:128-214,
    generalised path) from the adjoint of the converged fixed point:
    Omega_cc(beta, mu) with T = Phi(T) has dOmega_cc/dtheta =
    dE/dtheta + sum_y <Lam_y, dPhi_y/dtheta> where Lam solves
    Lam = dE/dT + (dPhi/dT)^T Lam (the lambda equations of
    kelvin/cc_utils.py:414-480, here in kelvin's L convention and converted:
    Lam1_y = -(g_y/beta) L1_y^T, Lam2_y = -(g_y/(4 beta)) L2_y^T).  The
    residual is linear in the integrals, so dPhi/dtheta through them is the
    integrated residual with the derivative blocks of response_blocks; the
    grid scales with beta (d_ft_quad, kelvin/ccsd.py:1075-1148).  First
    order: sys.g_mp1_den (kelvin/ccsd.py:1434-1445, :186-190)."""
    tconv = tconv if tconv is not None else 1000.0 * econv
    r = ft_ccsd(sys, T, mu, ngrid=ngrid, damp=damp, max_iter=max_iter,
                econv=econv, tconv=tconv, quad=quad, athresh=athresh)
    beta = 1.0 / T
    ti, g, G = ft_quad(ngrid, beta, quad)
    en = sys.g_energies_tot()
    n = en.shape[0]
    fo, fv = ff(beta, en, mu), ffv(beta, en, mu)
    fock, eri = sys.g_fock_tot(), sys.g_aint_tot()
    if athresh > 0.0:
        iocc = [i for i, x in enumerate(fo) if x > athresh]
        ivir = [i for i, x in enumerate(fv) if x > athresh]
    else:
        iocc = ivir = list(range(n))
    F, I = ft_active_integrals(fock, eri, en, fo[iocc], fv[ivir], iocc, ivir)
    d1 = D1(en, en)[numpy.ix_(ivir, iocc)]
    d2 = D2(en, en)[numpy.ix_(ivir, ivir, iocc, iocc)]
    T1, T2, ng = r["T1"], r["T2"], ngrid
    conv = {"tconv": tconv, "max_iter": max_iter, "damp": damp}
    L1, L2 = ccsd_lambda_guess(F, I, T1, beta, ng)
    L1, L2, lconv, lit = ft_lambda_iter(L1, L2, T1, T2, F, I, d1, d2, g, G,
                                        beta, ng, ti, conv)
    Lam1 = -(g / beta)[:, None, None] * L1.transpose(0, 2, 1)
    Lam2 = (-(g / (4.0 * beta))[:, None, None, None, None]
            * L2.transpose(0, 3, 4, 1, 2))
    Lt1 = int_adjoint(ng, Lam1, ti, d1, G)
    Lt2 = int_adjoint(ng, Lam2, ti, d2, G)
    Tb1, Tb2 = residual_tbar(F, I, T1, T2)
    Fd = sys.g_fock_d_den()                     # [p,q,k] = <pk||qk> fo fv

    def h_part(dx):
        dfock = -einsum('pqk,k->pq', Fd, dx)
        dF, dI = response_blocks(F, I, dfock, fo, fv, dx, iocc, ivir)
        out = ft_cc_energy(T1, T2, dF.ov, dI.oovv, g, beta)
        for y in range(ng):
            r1, r2 = stanton_rhs(dF, dI, T1[y], T2[y])
            out -= numpy.sum(Lt1[y] * r1) + numpy.sum(Lt2[y] * r2)
        return out

    dmu = h_part(-beta * numpy.ones(n))
    dbeta = h_part(en - mu) - r["Omega_cc"] / beta
    if gderiv:
        gd, Gd = d_ft_quad(ng, beta, quad)
        dbeta += ft_cc_energy(T1, T2, F.ov, I.oovv, gd, beta)
        dbeta += grid_derivative(Lam1, Lam2, Tb1, Tb2, ti, G, Gd, d1, d2,
                                 beta)
    else:                                   # kelvin/ccsd.py:1262-1288
        e_last = numpy.zeros(ng)
        e_last[-1] = 1.0
        dbeta += ft_cc_energy(T1, T2, F.ov, I.oovv, e_last, beta)
    Ncc = -dmu
    Ecc = beta * dbeta + mu * Ncc + r["Omega_cc"]
    # zero and first order (kelvin/ccsd.py:172-190)
    G0, G1 = r["G0"], r["G1"]
    N0 = fo.sum()
    E0 = beta * dGP0(beta, en, mu).sum() + mu * N0 + G0
    ron1 = sys.g_mp1_den()
    N1 = ron1.sum()
    B1 = numpy.dot((en - mu) / beta, ron1)
    E1 = beta * B1 + mu * N1 + G1
    N = N0 + N1 + Ncc
    E = E0 + E1 + Ecc
    Om = r["Omega"]
    S = -beta * (Om - E + mu * N)
    return dict(E=E, S=S, N=N, E0=E0, E1=E1, Ecc=Ecc, N0=N0, N1=N1,
                Ncc=Ncc, Omega=Om, Omega_cc=r["Omega_cc"], L1=L1, L2=L2,
                lambda_converged=lconv, lambda_iter=lit, T1=T1, T2=T2)


# ---------------------------------------------------------------------------
# One-particle response densities: kelvin/ft_cc_equations.py:704-722
# (ccsd_1rdm).  The per-point densities come from cqcpy's ccsd_1rdm_*_opt
# (absent); they are restated here as the derivatives of kelvin's
# Lagrangian with respect to the Fock blocks, which is what the reference's
# own tests check them against (kelvin/tests/test_ft_ccsd_rdm.py:12-170:
# (1/beta) sum p F = the Lagrangian with only that Fock block non-zero).
# ---------------------------------------------------------------------------
def ft_lagrangian(F, I, T1, T2, L1, L2, D1, D2, ti, ng, g, G, beta):
    """kelvin/tests/test_ft_ccsd_rdm.py:12-22 compute_ref (with the Stanton
    form of the residual in place of ccsd_simple, the same map)."""
    T1t, T2t = ccsd_stanton(F, I, T1, T2, D1, D2, ti, ng, G)
    E = ft_cc_energy(T1, T2, F.ov, I.oovv, g, beta)
    A1 = einsum('via,vai->v', L1, T1t) / beta
    A2 = 0.25 * einsum('vijab,vabij->v', L2, T2t) / beta
    return E - numpy.dot(A1, g) - numpy.dot(A2, g)


def ccsd_1rdm(T1, T2, L1, L2, D1, D2, ti, ng, g, G):
    """(pia, pba, pji, pai) of kelvin/ft_cc_equations.py:704-722: with the
    reverse-integrated l = int_L(L), pia = sum_x g_x l1 and the rest the
    derivatives of the per-point Lagrangian with respect to F.vv, F.oo and
    F.ov (the last with the energy's t1 term)."""
    l1 = int_L1(ng, L1, ti, D1, g, G)
    l2 = int_L2(ng, L2, ti, D2, g, G)
    pia = einsum('x,xia->ia', g, l1)
    dvv = (einsum('xia,xei->xae', l1, T1)
           + 0.5 * einsum('xijca,xceij->xae', l2, T2))
    pba = einsum('x,xab->ba', g, dvv)
    doo = (-einsum('xja,xam->xmj', l1, T1)
           - 0.5 * einsum('xkjab,xabkm->xmj', l2, T2))
    pji = einsum('x,xij->ji', g, doo)
    dov = (einsum('xem->xme', T1)
           - einsum('xia,xam,xei->xme', l1, T1, T1)
           + einsum('xia,xaeim->xme', l1, T2)
           - 0.5 * einsum('xijab,xaeij,xbm->xme', l2, T2, T1)
           - 0.5 * einsum('xijab,xabim,xej->xme', l2, T2, T1))
    pai = einsum('x,xia->ai', g, dov)
    return pia, pba, pji, pai


def normal_1rdm(pia, pba, pji, pai, fo, fv, beta):
    """kelvin/ccsd.py:1373-1387 (no active space): the occupation-scaled
    blocks over beta, and kelvin/ccsd.py:1995-2006: plus diag(fo)."""
    sfo, sfv = numpy.sqrt(fo), numpy.sqrt(fv)
    n1 = (pia * sfo[:, None] * sfv[None, :] + pba * sfv[:, None] * sfv[None, :]
          + pji * sfo[:, None] * sfo[None, :]
          + pai * sfv[:, None] * sfo[None, :]) / beta
    return n1 + numpy.diag(fo)


# ---------------------------------------------------------------------------
# Grand potential: kelvin/ft_cc_energy.py:7-32
# ---------------------------------------------------------------------------
def ft_cc_energy(T1, T2, f, eri, g, beta, Qterm=True):
    t2_temp = 0.25 * T2
    if Qterm:
        t2_temp = t2_temp + 0.5 * einsum('yai,ybj->yabij', T1, T1)
    t1_n = einsum('y,yai->ai', g, T1)
    t2_n = einsum('y,yabij->abij', g, t2_temp)
    Es1 = einsum('ai,ia->', t1_n, f)
    Es2 = einsum('abij,ijab->', t2_n, eri)
    return (Es1 + Es2) / beta


# ---------------------------------------------------------------------------
# Damped fixed point: kelvin/cc_utils.py:111-173
# ---------------------------------------------------------------------------
def ft_cc_iter(T1old, T2old, F, I, D1, D2, g, G, beta, ng, ti, conv,
               history=None):
    """Return (E, T1, T2, converged, niter); history collects (E, res)."""
    ethresh, tthresh = conv["econv"], conv["tconv"]
    max_iter, alpha = conv["max_iter"], conv["damp"]
    converged = False
    i = 0
    Eold = 888888888.888888888
    nl1 = numpy.linalg.norm(T1old) + 0.1
    nl2 = numpy.linalg.norm(T2old) + 0.1
    while i < max_iter and not converged:
        T1, T2 = ccsd_stanton(F, I, T1old, T2old, D1, D2, ti, ng, G)
        res1 = numpy.linalg.norm(T1 - T1old) / nl1
        res2 = numpy.linalg.norm(T2 - T2old) / nl2
        T1old = alpha * T1old + (1.0 - alpha) * T1
        T2old = alpha * T2old + (1.0 - alpha) * T2
        nl1 = numpy.linalg.norm(T1old) + 0.1
        nl2 = numpy.linalg.norm(T2old) + 0.1
        E = ft_cc_energy(T1old, T2old, F.ov, I.oovv, g, beta)
        if history is not None:
            history.append((E, res1 + res2))
        logging.info(' %2d  %.10f   %.4E' % (i + 1, E, res1 + res2))
        i += 1
        if numpy.abs(E - Eold) < ethresh and res1 + res2 < tthresh:
            converged = True
        Eold = E
    return Eold, T1old, T2old, converged, i


def int_tbar_single(ng, ig, tbar, ti, D, G):
    """kelvin/quadrature.py:348-369 (int_tbar1_single / int_tbar2_single):
    the causal integral at point ig alone, dt = t_x - t_ig for x < ig."""
    dt = numpy.zeros(ng)
    for x in range(ig):
        dt[x] = ti[x] - ti[ig]
    shp = (ng,) + (1,) * D.ndim
    gtemp = numpy.exp(dt.reshape(shp) * D[None])
    return numpy.einsum('x,x...->...', G[ig], gtemp * tbar)


def ft_cc_iter_extrap(F, I, D1, D2, g, G, beta, ng, ti, conv):
    """kelvin/cc_utils.py:176-242: converge the grid points one at a time.
    Point 0 is never iterated; point 1 starts from the integrated driving
    terms, later points from a linear extrapolation of the previous two;
    each point iterates residual -> single-point integral -> damping until
    the RMS change of T1 plus that of T2 is below tconv."""
    thresh, max_iter, alpha = conv["tconv"], conv["max_iter"], conv["damp"]
    nv, no = F.vo.shape
    t1bar = numpy.zeros((ng, nv, no))
    t2bar = numpy.zeros((ng, nv, nv, no, no))
    T1new = numpy.zeros((ng, nv, no))
    T2new = numpy.zeros((ng, nv, nv, no, no))
    for ig in range(ng):
        if ig == 0:
            t1bar[0] = -F.vo
            t2bar[0] = -I.vvoo
            continue
        elif ig == 1:
            t1bar[ig] = -F.vo
            t2bar[ig] = -I.vvoo
            T1new[ig] = int_tbar_single(ng, ig, t1bar, ti, D1, G)
            T2new[ig] = int_tbar_single(ng, ig, t2bar, ti, D2, G)
        else:
            c = (ti[ig] - ti[ig - 1]) / (ti[ig - 2] - ti[ig - 1])
            T1new[ig] = T1new[ig - 1] + (T1new[ig - 2] - T1new[ig - 1]) * c
            T2new[ig] = T2new[ig - 1] + (T2new[ig - 2] - T2new[ig - 1]) * c
        nl1 = numpy.sqrt(float(T1new[ig].size))
        nl2 = numpy.sqrt(float(T2new[ig].size))
        i, converged = 0, False
        while i < max_iter and not converged:
            r1, r2 = stanton_rhs(F, I, T1new[ig], T2new[ig])
            t1bar[ig] = -r1            # stanton_rhs includes f_ai, <ab||ij>
            t2bar[ig] = -r2
            T1 = int_tbar_single(ng, ig, t1bar, ti, D1, G)
            T2 = int_tbar_single(ng, ig, t2bar, ti, D2, G)
            res1 = numpy.linalg.norm(T1 - T1new[ig]) / nl1
            res2 = numpy.linalg.norm(T2 - T2new[ig]) / nl2
            T1new[ig] = alpha * T1new[ig] + (1.0 - alpha) * T1
            T2new[ig] = alpha * T2new[ig] + (1.0 - alpha) * T2
            i += 1
            converged = res1 + res2 < thresh
    return T1new, T2new


def ft_ccsd(sys, T, mu, ngrid=10, damp=0.0, max_iter=40, econv=1e-8,
            tconv=None, quad='lin', athresh=0.0, history=None,
            rt_iter="all"):
    """kelvin/ccsd.py:43-76 (construction) + :595-712 (_ft_ccsd, 'g' path;
    rt_iter='point' takes the pointwise schedule of :697-700).

    Returns a dict with Omega (total), Omega_cc, G0, G1, mp2, T1, T2, niter.
    """
    tconv = tconv if tconv is not None else 1000.0 * econv       # ccsd.py:54
    beta = 1.0 / T
    ti, g, G = ft_quad(ngrid, beta, quad)
    ng = ngrid
    en = sys.g_energies_tot()
    fo = ff(beta, en, mu)
    fv = ffv(beta, en, mu)
    n = en.shape[0]
    G0 = GP0(beta, en, mu).sum() + sys.const_energy()            # :636-638
    G1 = sys.get_mp1()                                            # :639
    fock, eri = sys.g_fock_tot(), sys.g_aint_tot()
    if athresh > 0.0:                                             # :642-660
        iocc = [i for i, x in enumerate(fo) if x > athresh]
        ivir = [i for i, x in enumerate(fv) if x > athresh]
        F, I = ft_active_integrals(fock, eri, en, fo[iocc], fv[ivir],
                                   iocc, ivir)
        d1, d2 = D1(en, en), D2(en, en)
        d1 = d1[numpy.ix_(ivir, iocc)]
        d2 = d2[numpy.ix_(ivir, ivir, iocc, iocc)]
    else:
        F, I = ft_integrals(fock, eri, en, beta, mu)
        d1, d2 = D1(en, en), D2(en, en)
    nv, no = F.vo.shape
    if rt_iter[0] != 'a':                                         # :697-700
        conv = {"econv": econv, "tconv": tconv, "max_iter": max_iter,
                "damp": damp}
        T1, T2 = ft_cc_iter_extrap(F, I, d1, d2, g, G, beta, ng, ti, conv)
        Ecc = ft_cc_energy(T1, T2, F.ov, I.oovv, g, beta)
        return dict(Omega=Ecc + G0 + G1, Omega_cc=Ecc, G0=G0, G1=G1,
                    mp2=None, T1=T1, T2=T2, niter=None, converged=None, n=n)
    # MP2-level guess: kelvin/ccsd.py:680-688
    T1old = -numpy.broadcast_to(F.vo, (ng, nv, no)).copy()
    T2old = -numpy.broadcast_to(I.vvoo, (ng, nv, nv, no, no)).copy()
    T1old = int_tbar1(ng, T1old, ti, d1, G)
    T2old = int_tbar2(ng, T2old, ti, d2, G)
    E2 = ft_cc_energy(T1old, T2old, F.ov, I.oovv, g, beta, Qterm=False)
    conv = {"econv": econv, "tconv": tconv, "max_iter": max_iter,
            "damp": damp}
    Ecc, T1, T2, converged, niter = ft_cc_iter(
        T1old, T2old, F, I, d1, d2, g, G, beta, ng, ti, conv, history)
    return dict(Omega=Ecc + G0 + G1, Omega_cc=Ecc, G0=G0, G1=G1, mp2=E2,
                T1=T1, T2=T2, niter=niter, converged=converged, n=n)


# ---------------------------------------------------------------------------
# Systems (the reference's integral providers).
# ---------------------------------------------------------------------------
class UEGBasis:
    """kelvin/ueg_utils.py:21-54 (literal loops, same stable ordering)."""

    def __init__(self, L, cutoff, norb=None):
        self.L = L
        kmax = numpy.sqrt(2 * cutoff)
        imax = int(numpy.ceil(L * kmax / (2 * numpy.pi)) + 0.1)
        basis = []
        for i in range(-imax, imax):
            for j in range(-imax, imax):
                for k in range(-imax, imax):
                    k2 = 4.0 * numpy.pi * numpy.pi / (L * L) * (i*i + j*j + k*k)
                    if k2 < 2.0 * cutoff:
                        basis.append((i, j, k))
        basis.sort(key=lambda x: x[0]*x[0] + x[1]*x[1] + x[2]*x[2])
        if norb is not None:
            basis = basis[:norb]
        self.basis = basis
        c = 2.0 * numpy.pi / L
        self.ks = [numpy.array((c*x[0], c*x[1], c*x[2])) for x in basis]
        self.Es = [0.5 * float(k @ k) for k in self.ks]

    def get_nbsf(self):
        return len(self.basis)

    def build_g2e_matrix(self, anti=True):
        """kelvin/ueg_utils.py:109-138, iterating over (p,q,r) and solving
        momentum conservation for s instead of testing all s."""
        nb = self.get_nbsf()
        n = 2 * nb
        lookup = {tuple(x): i for i, x in enumerate(self.basis)}
        aaa = 4.0 * numpy.pi / self.L ** 3
        c2 = (2.0 * numpy.pi / self.L) ** 2
        V = numpy.zeros((n, n, n, n))
        for p in range(n):
            xp, sp = self.basis[p % nb], p // nb
            for q in range(n):
                xq, sq = self.basis[q % nb], q // nb
                for r in range(n):
                    xr, sr = self.basis[r % nb], r // nb
                    if p == r or sp != sr:
                        continue
                    xs = tuple(xp[d] + xq[d] - xr[d] for d in range(3))
                    s0 = lookup.get(xs)
                    if s0 is None:
                        continue
                    s = s0 + nb * sq
                    if q == s:
                        continue
                    dk2 = c2 * sum((xp[d] - xr[d]) ** 2 for d in range(3))
                    V[p, q, r, s] = aaa / dk2
        if anti:
            return V - V.transpose(0, 1, 3, 2)
        return V


class UEGSystem:
    """kelvin/ueg_system.py:31-65 and the generalised ('g') methods used on
    the hot path: g_energies_tot :266-267, g_fock_tot :400-419,
    g_aint_tot :650-651, get_mp1 :126-131, const_energy :86-90."""

    def __init__(self, T, L, Emax, mu, norb=None):
        self.T, self.L, self.mu = T, L, mu
        self.basis = UEGBasis(L, Emax, norb=norb)
        beta = 1.0 / T if T > 0.0 else 1.0e20
        self.N = ff(beta, self.g_energies_tot(), mu).sum()
        self.den = self.N / L ** 3
        self.rs = (3 / (4.0 * numpy.pi * self.den)) ** (1.0 / 3.0)
        self.Ef = 0.5 * (3.0 * numpy.pi ** 2 * self.den) ** (2.0 / 3.0)
        self.Tf = T / self.Ef
        self._eri = None

    def verify(self, T, mu):
        return T == self.T and mu == self.mu

    def has_u(self):
        return False

    def const_energy(self):
        return 0.0

    def g_energies_tot(self):
        e = numpy.asarray(self.basis.Es)
        return numpy.hstack((e, e))

    def g_aint_tot(self):
        if self._eri is None:
            self._eri = self.basis.build_g2e_matrix()
        return self._eri

    def g_fock_tot(self):
        d = self.g_energies_tot()
        fo = ff(1.0 / self.T, d, self.mu)
        return numpy.diag(d) + einsum('prqr,r->pq', self.g_aint_tot(), fo)

    def get_mp1(self):
        fo = ff(1.0 / self.T, self.g_energies_tot(), self.mu)
        return 0.5 * einsum('ijij,i,j->', self.g_aint_tot(), fo, fo)

    def g_mp1_den(self):
        """kelvin/ueg_system.py:204-215: beta dE1/dx_p with x = beta (e - mu)."""
        beta = 1.0 / self.T
        en = self.g_energies_tot()
        fo, fv = ff(beta, en, self.mu), ffv(beta, en, self.mu)
        return -beta * einsum('ijij,i,j->i', self.g_aint_tot(), fo * fv, fo)

    def g_fock_d_den(self):
        """kelvin/ueg_system.py:498-510: JK[p,q,k] = <pk||qk> fo_k fv_k."""
        beta = 1.0 / self.T
        en = self.g_energies_tot()
        fo, fv = ff(beta, en, self.mu), ffv(beta, en, self.mu)
        return einsum('piqi,i->pqi', self.g_aint_tot(), fo * fv)


class PUEGSystem(UEGSystem):
    """kelvin/pueg_system.py:10-203 at finite temperature: the spin-polarised
    gas, one spin, so the orbitals are the plane waves themselves (energies
    :114-115 over kelvin/ueg_utils.py:73-74) and <pq||rs> = <pq|rs> - <pq|sr>
    of the spatial integrals (:197-203, the same-spin block of the
    unsymmetrised spin-orbital matrix); the Fock matrix (:141-168), E1
    (:71-80) and derivative densities (:95-103, :178-195) take UEGSystem's
    forms over these."""

    def g_energies_tot(self):
        return numpy.asarray(self.basis.Es)

    def g_aint_tot(self):
        if self._eri is None:
            nb = self.basis.get_nbsf()
            V = self.basis.build_g2e_matrix(anti=False)[:nb, :nb, :nb, :nb]
            self._eri = V - V.transpose(0, 1, 3, 2)
        return self._eri


class DFModelSystem:
    """Molecule-like model with density-fitted ERIs.

    norb spatial orbitals, energies ~U[-1,1] (sorted) drawn first, then
    B[Q,p,q] (Q=2*norb) with upper triangle ~N(0, sigma^2), mirrored.
    (pq|rs) = sum_Q B_Qpq B_Qrs; spin-blocked (alpha block, then beta) and
    antisymmetrised <pq||rs> = (pr|qs) d_sp,sr d_sq,ss - (ps|qr) d_sp,ss d_sq,sr.
    The Fock matrix is diag(energies), so every Fermi-weighted Fock block is
    zero; E1 = 1/2 sum_ij <ij||ij> fo_i fo_j (the UEG-'g' convention).
    """

    def __init__(self, T, mu, norb=8, seed=0, sigma=0.03):
        rng = numpy.random.default_rng(seed)
        self.T, self.mu = T, mu
        e = numpy.sort(rng.uniform(-1.0, 1.0, norb))
        B = numpy.zeros((2 * norb, norb, norb))
        iu = numpy.triu_indices(norb)
        for Q in range(2 * norb):
            B[Q][iu] = rng.normal(0.0, sigma, len(iu[0]))
            B[Q] = B[Q] + numpy.triu(B[Q], 1).T
        self.e_spatial, self.B = e, B
        chem = einsum('Qpq,Qrs->pqrs', B, B)
        n = 2 * norb
        V = numpy.zeros((n, n, n, n))
        for s1 in range(2):
            for s2 in range(2):
                o1, o2 = s1 * norb, s2 * norb
                # <pq|rs> = (pr|qs) with p,r spin s1 and q,s spin s2
                V[o1:o1+norb, o2:o2+norb, o1:o1+norb, o2:o2+norb] = \
                    chem.transpose(0, 2, 1, 3)
        self._eri = V - V.transpose(0, 1, 3, 2)
        self._en = numpy.hstack((e, e))

    def verify(self, T, mu):
        return T == self.T and mu == self.mu

    def has_u(self):
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
        fo = ff(1.0 / self.T, self._en, self.mu)
        return 0.5 * einsum('ijij,i,j->', self._eri, fo, fo)

    def g_mp1_den(self):
        """beta dE1/dx_p (E1 as get_mp1; kelvin/ueg_system.py:204-215)."""
        beta = 1.0 / self.T
        fo, fv = ff(beta, self._en, self.mu), ffv(beta, self._en, self.mu)
        return -beta * einsum('ijij,i,j->i', self._eri, fo * fv, fo)

    def g_fock_d_den(self):
        """The Fock matrix is diag(e): no occupation dependence."""
        n = self._en.shape[0]
        return numpy.zeros((n, n, n))


class GenericSystem:
    """Explicit (h, <pq||rs>, reference energies) system for brute-force
    checks.  Fock = h + sum_r <pr||qr> fo_r; E1 = <H - H0>_0 =
    sum_p (h_pp - e_p) fo_p + 1/2 sum_pq <pq||pq> fo_p fo_q (the content of
    kelvin/scf_system.py:54-68 with ft_mp.mp1, kelvin/ft_mp.py:22-29)."""

    def __init__(self, T, mu, h, eri, en):
        self.T, self.mu, self.h, self._eri, self._en = T, mu, h, eri, en

    def verify(self, T, mu):
        return T == self.T and mu == self.mu

    def has_u(self):
        return False

    def const_energy(self):
        return 0.0

    def g_energies_tot(self):
        return self._en.copy()

    def _fo(self):
        return ff(1.0 / self.T, self._en, self.mu)

    def g_fock_tot(self):
        return self.h + einsum('prqr,r->pq', self._eri, self._fo())

    def g_aint_tot(self):
        return self._eri

    def get_mp1(self):
        fo = self._fo()
        return (numpy.dot(numpy.diag(self.h) - self._en, fo)
                + 0.5 * einsum('ijij,i,j->', self._eri, fo, fo))

    def g_mp1_den(self):
        """beta dE1/dx_p with x = beta (e - mu) (E1 as get_mp1)."""
        beta = 1.0 / self.T
        fo, fv = self._fo(), ffv(beta, self._en, self.mu)
        return -beta * fo * fv * ((numpy.diag(self.h) - self._en)
                                  + einsum('ijij,j->i', self._eri, fo))

    def g_fock_d_den(self):
        """[p,q,k] = <pk||qk> fo_k fv_k (Fock = h + sum_k <pk||qk> fo_k)."""
        beta = 1.0 / self.T
        fo, fv = self._fo(), ffv(beta, self._en, self.mu)
        return einsum('piqi,i->pqi', self._eri, fo * fv)


def ueg_params(nk, rs, N, t_over_tf):
    """Box length, temperature and Fermi energy for a UEG with N electrons at
    Wigner-Seitz radius rs (L from kelvin/tests/test_ueg.py:8-10, E_F from
    kelvin/ueg_system.py:57-62)."""
    L = rs * (4.0 * numpy.pi * N / 3.0) ** (1.0 / 3.0)
    Ef = 0.5 * (3.0 * numpy.pi ** 2 * N / L ** 3) ** (2.0 / 3.0)
    return L, t_over_tf * Ef, Ef


def solve_mu(en, beta, N, lo=None, hi=None, tol=1e-14):
    """Bisection for sum_p ff(beta, e_p, mu) = N (non-interacting)."""
    lo = en.min() - 50.0 / beta if lo is None else lo
    hi = en.max() + 50.0 / beta if hi is None else hi
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if ff(beta, en, mid).sum() > N:
            hi = mid
        else:
            lo = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)
