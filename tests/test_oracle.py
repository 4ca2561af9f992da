"""Pins the CPU oracle (oracle/ftcc.py) before anything is compared with it.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign,
a transposed index, a wrong weight) fails one of them:
  * the residual against a brute-force Fock-space projection of
    e^{-T} H e^{T} (random, non-Hermitian Hamiltonians, o != v);
  * the reference's golden grand potentials and per-iteration energies;
  * FT-CCSD = FT-FCI for two spin orbitals (kelvin/tests/test_ft_ccsd.py:26-47);
  * closed forms for the quadrature, Fermi factors and integrators.
"""
import json
import os

import numpy
import pytest

from oracle import ftcc
from oracle import fock_space as fs

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_goldens.json")))


def _blocks(f, eri, no):
    n = f.shape[0]
    sl = {'o': slice(0, no), 'v': slice(no, n)}
    F = ftcc.Blocks(oo=f[sl['o'], sl['o']], ov=f[sl['o'], sl['v']],
                    vo=f[sl['v'], sl['o']], vv=f[sl['v'], sl['v']])
    I = ftcc.Blocks({k: eri[sl[k[0]], sl[k[1]], sl[k[2]], sl[k[3]]]
                     for k in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo',
                               'oovv', 'vooo', 'ooov', 'oooo')})
    return F, I


def random_amplitudes(nv, no, rng, scale=0.2):
    t1 = rng.normal(size=(nv, no)) * scale
    t2 = rng.normal(size=(nv, nv, no, no)) * scale
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, t2


@pytest.mark.parametrize("no,nv,herm", [(2, 3, True), (3, 3, False),
                                        (2, 4, False), (3, 2, False)])
def test_stanton_matches_bruteforce_projection(no, nv, herm):
    rng = numpy.random.default_rng(10 * no + nv)
    n = no + nv
    f = rng.normal(size=(n, n)) * 0.3
    if herm:
        f = f + f.T
    eri = fs.random_antisym(n, rng, hermitian=herm) * 0.1
    t1, t2 = random_amplitudes(nv, no, rng)
    F, I = _blocks(f, eri, no)
    r1, r2 = ftcc.stanton_rhs(F, I, t1, t2)
    b1, b2 = fs.cc_projection(f, eri, t1, t2, no)
    assert numpy.abs(r1 - b1).max() < 1e-13
    assert numpy.abs(r2 - b2).max() < 1e-13


def test_planted_mistake_is_caught():
    """A sign flip in one Wmbej term must break the brute-force agreement."""
    rng = numpy.random.default_rng(7)
    no, nv = 2, 3
    n = no + nv
    f = rng.normal(size=(n, n)) * 0.3
    eri = fs.random_antisym(n, rng, hermitian=False) * 0.1
    t1, t2 = random_amplitudes(nv, no, rng)
    F, I = _blocks(f, eri, no)
    I_bad = ftcc.Blocks(I)
    I_bad['vovo'] = I.vovo * 1.001
    r1, r2 = ftcc.stanton_rhs(F, I_bad, t1, t2)
    b1, b2 = fs.cc_projection(f, eri, t1, t2, no)
    assert numpy.abs(r2 - b2).max() > 1e-6


def _ueg(g):
    s = g["system"]
    L = 2 * numpy.pi if s["L"] == "2*pi" else s["L"]
    return ftcc.UEGSystem(s["T"], L, s["Emax"], s["mu"], norb=s["norb"]), s


def test_golden_G1_omega_cc():
    g = GOLD["G1"]
    ueg, s = _ueg(g)
    r = ftcc.ft_ccsd(ueg, s["T"], s["mu"], **g["ccsd"])
    assert abs(r["Omega_cc"] - g["Omega_cc"]) < g["tol"]


def test_golden_G2_pueg_omega_cc():
    """kelvin/tests/test_ft_ccsd.py:24,151-165: the spin-polarised gas
    (one spin, 7 plane waves) through the oracle's solver.  The golden
    value is the energy of iteration 10, the first whose change is below
    econv, to all 12 printed digits: it was printed by a run that tested
    the energy alone.  The loop as it stands (kelvin/cc_utils.py:158, energy
    and amplitude tests) stops at iteration 12, 3.3e-9 away, inside the
    test's 1e-8; the iteration-10 energy pins the integrals and the
    residual to 1e-12."""
    g = GOLD["G2"]
    s = g["system"]
    pueg = ftcc.PUEGSystem(s["T"], 2 * numpy.pi, s["Emax"], s["mu"],
                           norb=s["norb"])
    assert pueg.g_energies_tot().shape == (s["norb"],)
    hist = []
    r = ftcc.ft_ccsd(pueg, s["T"], s["mu"], history=hist, **g["ccsd"])
    assert abs(r["Omega_cc"] - g["Omega_cc"]) < g["tol"]
    assert abs(hist[9][0] - hist[8][0]) < 1e-8
    assert abs(hist[9][0] - g["Omega_cc"]) < 1e-12
    assert r["niter"] == 12


def test_golden_G4_iterations():
    g = GOLD["G4"]
    ueg, s = _ueg(g)
    assert abs(ueg.N - g["N0"]) < 1e-10
    hist = []
    r = ftcc.ft_ccsd(ueg, s["T"], s["mu"], history=hist, **g["ccsd"])
    assert abs(r["mp2"] - g["mp2"]) < 6e-11
    for (E, _), ref in zip(hist, g["iter_energies"]):
        assert abs(E - ref) < 6e-11
    assert abs(r["Omega"] - g["Omega"]) < 6e-11
    assert abs(r["Omega_cc"] - g["Omega_cc"]) < 6e-12


def test_golden_G3_first_iterations():
    g = GOLD["G3"]
    ueg, s = _ueg(g)
    assert abs(ueg.N - g["N0"]) < 1e-10
    hist = []
    cc = dict(g["ccsd"], max_iter=2)
    r = ftcc.ft_ccsd(ueg, s["T"], s["mu"], history=hist, **cc)
    assert abs(r["mp2"] - g["mp2"]) < 6e-11
    for (E, _), ref in zip(hist, g["iter_energies"][:2]):
        assert abs(E - ref) < 6e-11


def test_ft_fci_exact_two_spin_orbitals():
    eps, U = -0.3, 0.8
    h = numpy.diag([eps, eps])
    eri = numpy.zeros((2, 2, 2, 2))
    eri[0, 1, 0, 1] = eri[1, 0, 1, 0] = U
    eri[0, 1, 1, 0] = eri[1, 0, 0, 1] = -U
    T, mu = 1.0, 0.1
    sysm = ftcc.GenericSystem(T, mu, h, eri, numpy.array([eps, eps]) + 0.2)
    r = ftcc.ft_ccsd(sysm, T, mu, ngrid=200, max_iter=60, econv=1e-12)
    assert abs(r["Omega"] - fs.ft_fci_omega(h, eri, 1.0 / T, mu)) < 1e-9


# --- quadrature ----------------------------------------------------------
@pytest.mark.parametrize("ng", [9, 10, 17])
def test_simpson_lin_exactness(ng):
    beta = 2.3
    ti, g, G = ftcc.ft_quad(ng, beta, 'lin')
    assert ti[0] == 0.0 and abs(ti[-1] - beta) < 1e-14
    deg = 3 if ng % 2 == 1 else 1
    for k in range(deg + 1):
        assert abs(g @ ti ** k - beta ** (k + 1) / (k + 1)) < 1e-12
    for y in range(1, ng):
        degy = 3 if y % 2 == 0 else 1
        for k in range(degy + 1):
            assert abs(G[y] @ ti ** k - ti[y] ** (k + 1) / (k + 1)) < 1e-12
    assert numpy.all(G[0] == 0.0)
    assert numpy.all(numpy.triu(G, 1) == 0.0)


@pytest.mark.parametrize("quad", ['lin', 'ln', 'sin', 'exp', 'quad', 'cub',
                                  'quar', 'L'])
def test_beta_derivative_rules(quad):
    ng, beta, h = 11, 1.7, 1e-5
    gd, Gd = ftcc.d_ft_quad(ng, beta, quad)
    _, gp, Gp = ftcc.ft_quad(ng, beta + h, quad)
    _, gm, Gm = ftcc.ft_quad(ng, beta - h, quad)
    assert numpy.abs((gp - gm) / (2 * h) - gd).max() < 1e-8
    assert numpy.abs((Gp - Gm) / (2 * h) - Gd).max() < 1e-8


def test_mapped_rules_integrate():
    beta = 1.9
    for quad in ('ln', 'sin', 'exp', 'quad', 'cub', 'quar', 'mid'):
        ti, g, G = ftcc.ft_quad(41, beta, quad)
        assert abs(g.sum() - beta) < 1e-5
        assert abs(g @ ti - beta ** 2 / 2) < 1e-4
    with pytest.raises(Exception):
        ftcc.ft_quad(10, 1.0, 'xyz')


# --- Fermi factors ---------------------------------------------------------
def test_fermi_closed_forms():
    e = numpy.linspace(-3, 3, 41)
    beta, mu = 37.0, 0.2
    fo, fv = ftcc.ff(beta, e, mu), ftcc.ffv(beta, e, mu)
    assert numpy.abs(fo + fv - 1.0).max() < 1e-15
    x = beta * (e - mu)
    ok = numpy.abs(x) < 30
    assert numpy.allclose(fo[ok], 1.0 / (1.0 + numpy.exp(x[ok])), rtol=1e-14)
    assert numpy.allclose(ftcc.GP0(beta, e, mu)[ok],
                          -numpy.log(1 + numpy.exp(-x[ok])) / beta, rtol=1e-13)
    tiny = ftcc.ff(500.0, numpy.array([1.0]), 0.0)[0]
    assert abs(tiny / numpy.exp(-500.0) - 1.0) < 1e-12
    tinyv = ftcc.ffv(500.0, numpy.array([-1.0]), 0.0)[0]
    assert abs(tinyv / numpy.exp(-500.0) - 1.0) < 1e-12
    h = 1e-6
    dnum = (ftcc.GP0(beta + h, e, mu) - ftcc.GP0(beta - h, e, mu)) / (2 * h)
    assert numpy.abs(dnum - ftcc.dGP0(beta, e, mu)).max() < 1e-8
    E0 = beta * ftcc.dGP0(beta, e, mu).sum() + mu * fo.sum() + ftcc.GP0(beta, e, mu).sum()
    assert abs(E0 - (e * fo).sum()) < 1e-12


# --- integrators -----------------------------------------------------------
def test_integrator_constant_source_closed_form():
    """int_0^tau c e^{-(tau-s)D} ds = c (1-e^{-tau D})/D; the error must fall
    as delta^3 (the trapezoid first panel of odd rows dominates)."""
    rng = numpy.random.default_rng(2)
    beta = 3.0
    e = rng.uniform(-1, 1, 3)
    d1 = ftcc.D1(e, e)
    d1[numpy.abs(d1) < 1e-3] = 0.3
    c = rng.normal(size=d1.shape)
    errs = []
    for ng in (41, 81):
        ti, g, G = ftcc.ft_quad(ng, beta, 'lin')
        tb = numpy.broadcast_to(c, (ng,) + c.shape).copy()
        t = ftcc.int_tbar1(ng, tb, ti, d1, G)
        exact = c[None] * (1 - numpy.exp(-ti[:, None, None] * d1[None])) / d1[None]
        errs.append(numpy.abs(t - exact).max() / numpy.abs(exact).max())
    assert errs[0] < 1e-4
    assert 6.0 < errs[0] / errs[1] < 10.0


def test_integrator_literal_equals_scan():
    rng = numpy.random.default_rng(3)
    ng, n = 7, 3
    e = rng.uniform(-1, 1, n)
    d2 = ftcc.D2(e, e)
    tb = rng.normal(size=(ng, n, n, n, n))
    for quad in ('lin', 'mid', 'L', 'sin'):
        ti, g, G = ftcc.ft_quad(ng, 2.0, quad)
        a = ftcc.int_tbar2(ng, tb, ti, d2, G)
        b = ftcc.int_tbar_scan(ng, tb, ti, d2, G)
        assert numpy.abs(a - b).max() < 1e-13


@pytest.mark.parametrize("quad,step", [('lin', 2), ('mid', 1), ('L', 1)])
def test_recurrence_equals_literal_integrator(quad, step):
    """The device integrates equally spaced grids whose G rows repeat the row
    `step` above them left of the last three columns (Simpson: 2; midpoint
    and 'L': 1) by T_y = r^s T_{y-s} + sum_{x=y-2..y} (G[y,x] - G[y-s,x])
    r^(y-x) Tbar_x.  Restated here in numpy, it must equal the literal
    causal sum, including for D < 0 (r > 1)."""
    rng = numpy.random.default_rng(5)
    ng, n = 24, 3
    e = rng.uniform(-1.5, 1.5, n)
    d2 = ftcc.D2(e, e)
    tb = rng.normal(size=(ng, n, n, n, n))
    ti, g, G = ftcc.ft_quad(ng, 8.0, quad)
    Gs = numpy.vstack([numpy.zeros((step, ng)), G])   # Gs[y + step] = G[y]
    for y in range(ng):                               # the structure relied on
        assert numpy.array_equal(G[y, :max(0, y - 2)], Gs[y, :max(0, y - 2)])
    assert not numpy.triu(G, 1).any()
    delta = ti[1] - ti[0]
    r = numpy.exp(-delta * d2)
    out = numpy.zeros_like(tb)
    for y in range(ng):
        t = r ** step * out[y - step] if y >= step else numpy.zeros_like(d2)
        for x in range(max(0, y - 2), y + 1):
            t = t + (G[y, x] - Gs[y, x]) * r ** (y - x) * tb[x]
        out[y] = t
    ref = ftcc.int_tbar2(ng, tb, ti, d2, G)
    assert numpy.abs(out - ref).max() < 1e-12 * numpy.abs(ref).max()
    if step == 2:          # the Simpson rows do not repeat the row just above
        assert not all(numpy.array_equal(G[y, :y - 2], G[y - 1, :y - 2])
                       for y in range(3, ng))


# --- systems ----------------------------------------------------------------
def test_ueg_integral_symmetries():
    ueg = ftcc.UEGSystem(0.1, 3.8855, 7.4, 2.0, norb=19)
    I = ueg.g_aint_tot()
    assert numpy.abs(I + I.transpose(1, 0, 2, 3)).max() < 1e-14
    assert numpy.abs(I + I.transpose(0, 1, 3, 2)).max() < 1e-14
    assert numpy.abs(I - I.transpose(2, 3, 0, 1)).max() < 1e-14


def test_df_model_system_shapes():
    s = ftcc.DFModelSystem(0.1, 0.0, norb=8, seed=0)
    eri = s.g_aint_tot()
    assert eri.shape == (16,) * 4
    assert numpy.abs(eri + eri.transpose(1, 0, 2, 3)).max() < 1e-15
    assert numpy.abs(eri - eri.transpose(2, 3, 0, 1)).max() < 1e-15


def test_pointwise_schedule_reaches_the_same_fixed_point():
    """kelvin/tests/test_ft_ccsd.py:167-184: rt_iter='point' and the
    all-at-once iteration give the same Omega_cc within 1e-8."""
    u = ftcc.UEGSystem(0.1, 2 * numpy.pi, 1.2, 0.1, norb=7)
    kw = dict(ngrid=10, damp=0.2, max_iter=50, tconv=1e-8)
    r1 = ftcc.ft_ccsd(u, 0.1, 0.1, **kw)
    r2 = ftcc.ft_ccsd(u, 0.1, 0.1, rt_iter='point', **kw)
    assert abs(r1["Omega_cc"] - r2["Omega_cc"]) < 1e-8
    assert abs(r2["Omega_cc"] - GOLD["G1"]["Omega_cc"]) < 1e-8


# ------------------------------------------------------------ TD-CCSD
def _linear_td_problem(no, nv, rng):
    """Interaction-free problem: only F.vo drives T1, so
    T1(tau) = -F.vo (1 - exp(-D1 tau)) / D1 exactly and T2 stays zero."""
    F = ftcc.Blocks(oo=numpy.zeros((no, no)), ov=numpy.zeros((no, nv)),
                    vo=rng.normal(size=(nv, no)), vv=numpy.zeros((nv, nv)))
    I = ftcc.Blocks({k: numpy.zeros([no if c == 'o' else nv for c in k])
                     for k in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo',
                               'oovv', 'vooo', 'ooov', 'oooo')})
    D1 = rng.uniform(0.5, 2.0, size=(nv, no))
    D2 = rng.uniform(0.5, 2.0, size=(nv, nv, no, no))
    return F, I, D1, D2


@pytest.mark.parametrize("order", [1, 2, 4])
def test_td_propagator_order_on_the_linear_problem(order):
    """kelvin/propagation.py rk1/rk2/rk4 have global order 1, 2 and 4: the
    error against the exact solution falls by 2^order when the step halves
    (a wrong stage coefficient drops the order)."""
    from oracle import td_ccsd
    rng = numpy.random.default_rng(order)
    no, nv = 2, 3
    F, I, D1, D2 = _linear_td_problem(no, nv, rng)
    tend = 2.0
    exact = -F.vo * (1.0 - numpy.exp(-D1 * tend)) / D1
    errs = []
    for nstep in (20, 40):
        ti = numpy.linspace(0.0, tend, nstep + 1)
        _, t1, t2 = td_ccsd.td_ccsd(F, I, D1, D2, ti, numpy.zeros(nstep + 1),
                                    1.0, order)
        assert numpy.abs(t2).max() == 0.0
        errs.append(numpy.abs(t1 - exact).max())
    ratio = errs[0] / errs[1]
    assert 0.8 * 2 ** order < ratio < 1.25 * 2 ** order, (errs, ratio)


@pytest.mark.parametrize("order,ng_td,ng_cc,tol", [(2, 60, 20, 1e-4),
                                                   (4, 60, 70, 1e-6)])
def test_td_ccsd_matches_fixed_point(order, ng_td, ng_cc, tol):
    """kelvin/tests/test_td_ccsd.py:33-75: the propagated Omega_cc agrees
    with the fixed-point FT-CCSD one (rk2 at 1e-4, rk4 at 1e-6), here on the
    7 k-point UEG of the reference's golden test."""
    from oracle import td_ccsd
    T, mu = 0.1, 0.1
    u = ftcc.UEGSystem(T, 2 * numpy.pi, 1.2, mu, norb=7)
    beta = 1.0 / T
    ref = ftcc.ft_ccsd(u, T, mu, ngrid=ng_cc, damp=0.2, max_iter=80,
                       econv=1e-11)
    en = u.g_energies_tot()
    F, I = ftcc.ft_integrals(u.g_fock_tot(), u.g_aint_tot(), en, beta, mu)
    D1, D2 = ftcc.D1(en, en), ftcc.D2(en, en)
    ti, g, _ = ftcc.ft_quad(ng_td, beta, 'lin')
    Ecc, _, _ = td_ccsd.td_ccsd(F, I, D1, D2, ti, g, beta, order)
    assert abs(Ecc - ref["Omega_cc"]) < tol, (Ecc, ref["Omega_cc"])


@pytest.mark.parametrize("scheme,order", [("be", 1), ("cn", 2), ("am2", 2)])
def test_td_implicit_propagator_order(scheme, order):
    """kelvin/propagation.py be/cn/am2 with the inner iteration converged
    tightly: error ratios 2, 4, 4 on halving the step of the linear problem."""
    from oracle import td_ccsd
    rng = numpy.random.default_rng(11 + order)
    no, nv = 2, 3
    F, I, D1, D2 = _linear_td_problem(no, nv, rng)
    tend = 2.0
    exact = -F.vo * (1.0 - numpy.exp(-D1 * tend)) / D1
    errs = []
    for nstep in (20, 40):
        ti = numpy.linspace(0.0, tend, nstep + 1)
        _, t1, _ = td_ccsd.td_ccsd(F, I, D1, D2, ti, numpy.zeros(nstep + 1),
                                   1.0, None,
                                   implicit=(scheme, 400, 0.0, 1e-13))
        errs.append(numpy.abs(t1 - exact).max())
    ratio = errs[0] / errs[1]
    assert 0.8 * 2 ** order < ratio < 1.25 * 2 ** order, (errs, ratio)


# ------------------------------------------------------------ C1 model
def test_df_model_integrals_pinned():
    """The density-fitted model's <pq||rs> carries (pq|rs) = sum_Q B_Qpq
    B_Qrs in the right slots: its mixed-spin block (p, r alpha; q, s beta)
    is <pq|rs> = (pr|qs) alone, which as a matrix over (p,r) x (q,s) must be
    positive semidefinite of rank at most 2*norb (it is B^T B), with the
    8-fold symmetry of real orbitals; the spin-orbital tensor vanishes unless
    the spins of (p, q) equal those of (r, s) as a set.  A (pq|rs)-for-
    (pr|qs) slip breaks the semidefiniteness and the rank."""
    norb = 8
    s = ftcc.DFModelSystem(0.1, 0.0, norb=norb, seed=0)
    eri = s.g_aint_tot()
    n = 2 * norb
    a, b = slice(0, norb), slice(norb, n)
    chem = eri[a, b, a, b].transpose(0, 2, 1, 3)     # (pr|qs) -> [p,r,q,s]
    assert numpy.abs(chem).max() > 1e-4
    M = chem.reshape(norb * norb, norb * norb)
    w = numpy.linalg.eigvalsh(0.5 * (M + M.T))
    assert w.min() > -1e-14 * w.max()
    assert int((w > 1e-12 * w.max()).sum()) <= 2 * norb
    assert numpy.abs(M - M.T).max() < 1e-16
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (1, 0, 3, 2)):
        assert numpy.abs(chem - chem.transpose(perm)).max() < 1e-16
    # the same B gives the same (pq|rs)
    ref = numpy.einsum('Qpq,Qrs->pqrs', s.B, s.B)
    assert numpy.abs(chem - ref).max() < 1e-16
    # spin pattern of the antisymmetrised integrals
    sp = numpy.arange(n) // norb
    allowed = ((sp[:, None, None, None] == sp[None, None, :, None]) &
               (sp[None, :, None, None] == sp[None, None, None, :])) | \
              ((sp[:, None, None, None] == sp[None, None, None, :]) &
               (sp[None, :, None, None] == sp[None, None, :, None]))
    assert numpy.abs(eri[~allowed]).max() == 0.0
    assert numpy.abs(eri[allowed]).max() > 1e-4
    # same-spin block: <pq||rs> = (pr|qs) - (ps|qr)
    same = chem.transpose(0, 2, 1, 3) - chem.transpose(0, 2, 3, 1)
    assert numpy.abs(eri[a, a, a, a] - same).max() < 1e-16


@pytest.mark.skipif(os.environ.get("FTCC_SLOW") != "1",
                    reason="several minutes of numpy at n = 38 "
                           "(FTCC_SLOW=1 runs it)")
def test_c3_corner_diverges_in_the_reference_algorithm():
    """At rs = 4, T = 0.05 T_F (the one point of the 33-k-point sweep where
    the GPU's damped iteration does not converge) the reference's own
    algorithm diverges too: the oracle at 19 k-points (n = 38), ng = 16,
    damping 0.6 runs away (|E| > 1e3 within 30 iterations; NaN at iteration
    28 when continued)."""
    L, T, Ef = ftcc.ueg_params(19, 4.0, 14, 0.05)
    Emax = 0.5 * (2 * numpy.pi / L) ** 2 * 8.0
    s0 = ftcc.UEGSystem(T, L, Emax, 0.0, norb=19)
    mu = ftcc.solve_mu(s0.g_energies_tot(), 1.0 / T, 14)
    s = ftcc.UEGSystem(T, L, Emax, mu, norb=19)

    class Stop(Exception):
        pass

    class Watch(list):
        def append(self, x):
            super().append(x)
            if not numpy.isfinite(x[0]) or abs(x[0]) > 1e3:
                raise Stop()

    hist = Watch()
    with numpy.errstate(all="ignore"):
        with pytest.raises(Stop):
            ftcc.ft_ccsd(s, T, mu, ngrid=16, damp=0.6, max_iter=30,
                         econv=1e-10, tconv=1e-7, history=hist)
    assert len(hist) <= 30


# ------------------------------------------------------------ adjoint
def _random_blocks(no, nv, rng):
    n = no + nv
    f = rng.normal(size=(n, n)) * 0.3
    eri = fs.random_antisym(n, rng, hermitian=False) * 0.1
    sl = {'o': slice(0, no), 'v': slice(no, n)}
    F = ftcc.Blocks(oo=f[sl['o'], sl['o']], ov=f[sl['o'], sl['v']],
                    vo=f[sl['v'], sl['o']], vv=f[sl['v'], sl['v']])
    I = ftcc.Blocks({k: numpy.ascontiguousarray(
        eri[sl[k[0]], sl[k[1]], sl[k[2]], sl[k[3]]])
        for k in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo', 'oovv', 'vooo',
                  'ooov', 'oooo')})
    return F, I


def _jt_mismatch(F, I, t1, t2, v1, v2, w1, w2, g1, g2, h=1.0):
    """|<g, w> - d/dh <v, stanton_rhs(t + h w)>| / |...|: the residual is a
    polynomial of degree 4 in h, so the five-point central difference is
    its derivative at 0 exactly (up to rounding), at any h."""
    def f(x):
        r1, r2 = ftcc.stanton_rhs(F, I, t1 + x * w1, t2 + x * w2)
        return (v1 * r1).sum() + (v2 * r2).sum()
    d = (f(-2 * h) - 8 * f(-h) + 8 * f(h) - f(2 * h)) / (12 * h)
    a = (g1 * w1).sum() + (g2 * w2).sum()
    return abs(a - d) / abs(d)


@pytest.mark.parametrize("no,nv,seed", [(3, 4, 3), (4, 2, 7), (2, 5, 11)])
def test_residual_transpose_is_exact(no, nv, seed):
    """stanton_rhs_transpose (the Lambda-equation operator) is the transpose
    of the residual's Jacobian: <J^T v, w> equals the derivative of
    <v, R(t + h w)> for random non-Hermitian blocks with o != v and
    unsymmetric v, w (every component independent), to rounding; a 1e-6
    change of one component of the result is detected."""
    rng = numpy.random.default_rng(seed)
    F, I = _random_blocks(no, nv, rng)
    t1 = rng.normal(size=(nv, no)) * 0.3
    t2 = rng.normal(size=(nv, nv, no, no)) * 0.3
    v1, v2 = rng.normal(size=t1.shape), rng.normal(size=t2.shape)
    w1, w2 = rng.normal(size=t1.shape), rng.normal(size=t2.shape)
    g1, g2 = ftcc.stanton_rhs_transpose(F, I, t1, t2, v1, v2)
    for h in (0.5, 1.0):
        assert _jt_mismatch(F, I, t1, t2, v1, v2, w1, w2, g1, g2, h) < 1e-12
    # sensitivity: a planted error far below any single term is caught
    g2b = g2.copy()
    g2b[0, 1, 1, 0] += 1e-6 * numpy.abs(g2).max()
    w2b = numpy.zeros_like(w2)
    w2b[0, 1, 1, 0] = 1.0
    assert _jt_mismatch(F, I, t1, t2, v1, v2, w1 * 0, w2b, g1, g2b) > 1e-8


def test_adjoint_integration_is_the_transpose():
    """int_adjoint is the transpose of int_tbar1/2 over the slices (uniform
    Simpson and the non-uniform 'sin' grid, any G), and kelvin's int_L1 /
    int_L2 (kelvin/quadrature.py:320-345) are it in the g-scaled convention:
    int_L(L)[s] = int_adjoint(g L)[s] / g[s]."""
    rng = numpy.random.default_rng(4)
    nv, no = 3, 2
    D1 = rng.uniform(-2, 2, size=(nv, no))
    D2 = rng.uniform(-2, 2, size=(nv, nv, no, no))
    for quad in ('lin', 'sin'):
        ng = 7
        ti, g, G = ftcc.ft_quad(ng, 2.0, quad)
        for D in (D1, D2):
            tb = rng.normal(size=(ng,) + D.shape)
            lam = rng.normal(size=(ng,) + D.shape)
            fwd = ftcc.int_tbar_scan(ng, tb, ti, D, G)
            adj = ftcc.int_adjoint(ng, lam, ti, D, G)
            a, b = (lam * fwd).sum(), (adj * tb).sum()
            assert abs(a - b) < 1e-13 * abs(a)
        L1 = rng.normal(size=(ng, no, nv))
        ref = ftcc.int_L1(ng, L1, ti, D1, g, G)
        got = ftcc.int_adjoint(ng, g[:, None, None] * L1.transpose(0, 2, 1),
                               ti, D1, G) / g[:, None, None]
        assert numpy.abs(got.transpose(0, 2, 1) - ref).max() < 1e-13
        L2 = rng.normal(size=(ng, no, no, nv, nv))
        ref = ftcc.int_L2(ng, L2, ti, D2, g, G)
        got = ftcc.int_adjoint(
            ng, g[:, None, None, None, None] * L2.transpose(0, 3, 4, 1, 2),
            ti, D2, G) / g[:, None, None, None, None]
        assert numpy.abs(got.transpose(0, 3, 4, 1, 2) - ref).max() < 1e-13


def test_esn_golden_G4():
    """E, S, N from the adjoint of the fixed point equal the values kelvin's
    lambda-equation compute_ESN printed for examples/ueg_ft_cc_compare.py
    (examples/ueg_ft_cc_compare.out:46-48; 'u' orbitals there, spin orbitals
    here), measured 9e-12, 9e-11 and 3e-11 off."""
    g = GOLD["G4"]
    s = g["system"]
    sysg = ftcc.UEGSystem(s["T"], 2 * numpy.pi, s["Emax"], s["mu"],
                          norb=s["norb"])
    r = ftcc.ft_ccsd_esn(sysg, s["T"], s["mu"], **g["ccsd"])
    assert r["lambda_converged"]
    for k in ("E", "S", "N"):
        assert abs(r[k] - g[k]) < 1e-9, (k, r[k], g[k])


def _omega(make, T, mu, kw):
    return ftcc.ft_ccsd(make(T, mu), T, mu, **kw)["Omega"]


@pytest.mark.parametrize("kind", ["df", "generic", "pueg"])
def test_esn_matches_finite_differences(kind):
    """N = -dOmega/dmu, S = -dOmega/dT and E = Omega + T S + mu N from the
    adjoint equal central differences (h = 1e-4) of fully converged solves
    on two small systems: the density-fitted model (E1 occupation
    dependent, Fock fixed) and a generic system whose Fock matrix depends
    on the occupations (kelvin's g_fock_d_den term)."""
    rng = numpy.random.default_rng(5)
    if kind == "df":
        def make(T, mu):
            return ftcc.DFModelSystem(T, mu, norb=3, seed=1, sigma=0.2)
        T, mu = 0.3, 0.05
    elif kind == "pueg":
        # kelvin/tests/test_ft_deriv.py:387-441 (the spin-polarised gas)
        def make(T, mu):
            return ftcc.PUEGSystem(T, 2 * numpy.pi, 1.2, mu, norb=7)
        T, mu = 0.1, 0.1
    else:
        n = 6
        h = rng.normal(size=(n, n)) * 0.2
        h = 0.5 * (h + h.T) + numpy.diag(numpy.linspace(-0.6, 0.6, n))
        eri = fs.random_antisym(n, rng, hermitian=True) * 0.08
        en = numpy.diag(h).copy() + 0.03

        def make(T, mu):
            return ftcc.GenericSystem(T, mu, h, eri, en)
        T, mu = 0.25, -0.02
    kw = dict(ngrid=8, damp=0.2, max_iter=200, econv=1e-13, tconv=1e-11)
    r = ftcc.ft_ccsd_esn(make(T, mu), T, mu, **kw)
    d = 1e-4
    N = -(_omega(make, T, mu + d, kw) - _omega(make, T, mu - d, kw)) / (2 * d)
    S = -(_omega(make, T + d, mu, kw) - _omega(make, T - d, mu, kw)) / (2 * d)
    E = r["Omega"] + T * S + mu * N
    assert abs(r["N"] - N) < 1e-7
    assert abs(r["S"] - S) < 1e-7
    assert abs(r["E"] - E) < 1e-7


def test_ccsd_1rdm_is_the_lagrangian_derivative():
    """The oracle's FT 1-RDM blocks (kelvin/ft_cc_equations.py:704-722)
    against the derivative of kelvin's Lagrangian with respect to each Fock
    element (kelvin/tests/test_ft_ccsd_rdm.py:12-170 checks the trace of
    the same identity), evaluated exactly by unit perturbations; random
    amplitudes, lambdas and denominators, ng = 5."""
    rng = numpy.random.default_rng(3)
    n, ng, beta = 4, 5, 2.0

    def antisym(x):
        x = x - numpy.swapaxes(x, -4, -3)
        return x - numpy.swapaxes(x, -2, -1)
    T1 = rng.normal(size=(ng, n, n)) * 0.3
    T2 = antisym(rng.normal(size=(ng, n, n, n, n)) * 0.2)
    L1 = rng.normal(size=(ng, n, n)) * 0.3
    L2 = antisym(rng.normal(size=(ng, n, n, n, n)) * 0.2)
    D1 = rng.normal(size=(n, n))
    D2 = rng.normal(size=(n, n, n, n))
    D2 = 0.5 * (D2 + D2.transpose(1, 0, 3, 2))
    ti, g, G = ftcc.ft_quad(ng, beta, 'lin')
    Z = numpy.zeros
    keys = ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo', 'oovv', 'vooo', 'ooov',
            'oooo')
    I = ftcc.Blocks({k: Z((n, n, n, n)) for k in keys})

    def brute(block):
        P = Z((n, n))
        for p in range(n):
            for q in range(n):
                F = ftcc.Blocks(oo=Z((n, n)), ov=Z((n, n)), vo=Z((n, n)),
                                vv=Z((n, n)))
                F[block][p, q] = 1.0
                P[q, p] = beta * ftcc.ft_lagrangian(F, I, T1, T2, L1, L2, D1,
                                                    D2, ti, ng, g, G, beta)
        return P
    got = ftcc.ccsd_1rdm(T1, T2, L1, L2, D1, D2, ti, ng, g, G)
    for block, p in zip(('vo', 'vv', 'oo', 'ov'), got):
        ref = brute(block)
        assert numpy.abs(p - ref).max() < 1e-12 * max(1.0, numpy.abs(ref).max())
    # a planted mistake (the l2 term of pba with the wrong sign) is caught
    bad = got[1] - 2 * 0.5 * numpy.einsum(
        'x,xab->ba', g, numpy.einsum('xijca,xceij->xae',
                                     ftcc.int_L2(ng, L2, ti, D2, g, G), T2))
    assert numpy.abs(bad - brute('vv')).max() > 1e-3


def test_lambda_makes_the_lagrangian_stationary():
    """kelvin/tests/test_ft_lambda.py:16-73 (fd_test_L1; :75-133 for the
    doubles): at the converged amplitudes and the lambdas of ft_lambda_iter,
    L(T) = E(T) - (1/beta) sum_y g_y [L1.(Phi1(T) - T1) + 1/4 L2.(Phi2(T) - T2)]
    has zero central-difference derivative (d = 1e-4) along every singles
    amplitude and along random antisymmetric doubles directions, on a
    generic system whose Fock matrix has a vo block (T1 is not zero); E
    alone does not, and lambdas off by 1% do not."""
    rng = numpy.random.default_rng(5)
    n, T, mu, ng = 6, 0.25, -0.02, 8
    h = rng.normal(size=(n, n)) * 0.2
    h = 0.5 * (h + h.T) + numpy.diag(numpy.linspace(-0.6, 0.6, n))
    eri = fs.random_antisym(n, rng, hermitian=True) * 0.08
    en = numpy.diag(h).copy() + 0.03
    sysx = ftcc.GenericSystem(T, mu, h, eri, en)
    r = ftcc.ft_ccsd(sysx, T, mu, ngrid=ng, damp=0.2, max_iter=300,
                     econv=1e-13, tconv=1e-11)
    assert r["converged"]
    beta = 1.0 / T
    ti, g, G = ftcc.ft_quad(ng, beta, 'lin')
    F, I = ftcc.ft_integrals(sysx.g_fock_tot(), sysx.g_aint_tot(), en, beta,
                             mu)
    D1, D2 = ftcc.D1(en, en), ftcc.D2(en, en)
    T1, T2 = r["T1"], r["T2"]
    assert numpy.abs(T1).max() > 1e-4
    L1, L2 = ftcc.ccsd_lambda_guess(F, I, T1, beta, ng)
    L1, L2, conv, _ = ftcc.ft_lambda_iter(
        L1, L2, T1, T2, F, I, D1, D2, g, G, beta, ng, ti,
        {"tconv": 1e-12, "max_iter": 300, "damp": 0.2})
    assert conv

    def lag(t1, t2, l1, l2):
        c = (numpy.einsum('yia,yai->y', l1, t1)
             + 0.25 * numpy.einsum('yijab,yabij->y', l2, t2)) / beta
        return (ftcc.ft_lagrangian(F, I, t1, t2, l1, l2, D1, D2, ti, ng, g,
                                   G, beta) + numpy.dot(c, g))

    def energy(t1, t2):
        return ftcc.ft_cc_energy(t1, t2, F.ov, I.oovv, g, beta)

    d = 1e-4

    def slope(f, x1, x2):
        return (f(T1 + d * x1, T2 + d * x2)
                - f(T1 - d * x1, T2 - d * x2)) / (2 * d)
    worst, e_scale = 0.0, 0.0
    for y in range(ng):
        for a in range(n):
            for i in range(n):
                x1 = numpy.zeros_like(T1)
                x1[y, a, i] = 1.0
                x2 = numpy.zeros_like(T2)
                worst = max(worst, abs(slope(
                    lambda t1, t2: lag(t1, t2, L1, L2), x1, x2)))
                e_scale = max(e_scale, abs(slope(energy, x1, x2)))
    for _ in range(3):
        x2 = ftcc.antisymmetrize(rng.normal(size=T2.shape))
        x1 = numpy.zeros_like(T1)
        worst = max(worst, abs(slope(lambda t1, t2: lag(t1, t2, L1, L2),
                                     x1, x2)))
        e_scale = max(e_scale, abs(slope(energy, x1, x2)))
    assert e_scale > 1e-3
    assert worst < 1e-7 * e_scale
    # lambdas off by 1%: the derivative no longer vanishes
    x1 = numpy.zeros_like(T1)
    x1[0, 0, 0] = 1.0
    off = abs(slope(lambda t1, t2: lag(t1, t2, 1.01 * L1, 1.01 * L2), x1,
                    numpy.zeros_like(T2)))
    assert off > 1e-5 * e_scale


def test_normal_1rdm_is_the_orbital_fock_derivative():
    """normal_1rdm (kelvin/ccsd.py:1373-1387) without its occupations is
    the derivative of the Lagrangian with respect to the orbital Fock
    matrix: with F[p,q] = s_p f[p,q] s_q (ft_integrals, kelvin/
    cc_utils.py:569-602, s the square roots of the occupations or
    vacancies), d L / d f[p,q] = n1[q,p], evaluated exactly by unit
    perturbations of f through ft_integrals; random amplitudes, multipliers
    and integrals, non-uniform occupations.  The occupation factors swapped
    in one block are detected."""
    rng = numpy.random.default_rng(8)
    n, ng, beta, mu = 4, 5, 2.0, 0.1
    en = numpy.sort(rng.uniform(-1.0, 1.0, n))

    def antisym(x):
        x = x - numpy.swapaxes(x, -4, -3)
        return x - numpy.swapaxes(x, -2, -1)
    T1 = rng.normal(size=(ng, n, n)) * 0.3
    T2 = antisym(rng.normal(size=(ng, n, n, n, n)) * 0.2)
    L1 = rng.normal(size=(ng, n, n)) * 0.3
    L2 = antisym(rng.normal(size=(ng, n, n, n, n)) * 0.2)
    eri = fs.random_antisym(n, rng, hermitian=True) * 0.1
    f0 = rng.normal(size=(n, n)) * 0.2 + numpy.diag(en)
    ti, g, G = ftcc.ft_quad(ng, beta, 'lin')
    D1, D2 = ftcc.D1(en, en), ftcc.D2(en, en)

    def lag(fock):
        F, I = ftcc.ft_integrals(fock, eri, en, beta, mu)
        return ftcc.ft_lagrangian(F, I, T1, T2, L1, L2, D1, D2, ti, ng, g,
                                  G, beta)
    base = lag(f0)
    deriv = numpy.zeros((n, n))
    for p in range(n):
        for q in range(n):
            f = f0.copy()
            f[p, q] += 1.0
            deriv[p, q] = lag(f) - base     # exact: L is linear in f
    fo, fv = ftcc.ff(beta, en, mu), ftcc.ffv(beta, en, mu)
    blocks = ftcc.ccsd_1rdm(T1, T2, L1, L2, D1, D2, ti, ng, g, G)
    n1 = ftcc.normal_1rdm(*blocks, fo, fv, beta) - numpy.diag(fo)
    scale = numpy.abs(deriv).max()
    assert scale > 1e-3
    assert numpy.abs(n1.T - deriv).max() < 1e-12 * max(1.0, scale)
    # the vo block scaled by sfv x sfo instead of sfo x sfv is caught
    pia, pba, pji, pai = blocks
    sfo, sfv = numpy.sqrt(fo), numpy.sqrt(fv)
    bad = n1 + (pia * sfv[:, None] * sfo[None, :]
                - pia * sfo[:, None] * sfv[None, :]) / beta
    assert numpy.abs(bad.T - deriv).max() > 1e-4 * scale


def test_active_integrals_are_the_full_blocks_restricted():
    """ft_active_integrals (kelvin/cc_utils.py:825-858) with the occupied
    list of orbitals whose occupation exceeds athresh and the virtual list
    of those whose vacancy does equals ft_integrals' blocks restricted to
    those lists, every one of the 13 blocks; and a solve whose threshold
    keeps every orbital (the active branch of kelvin/ccsd.py:642-660) equals
    the full-space solve."""
    rng = numpy.random.default_rng(12)
    n, T, mu = 6, 0.25, -0.02
    h = rng.normal(size=(n, n)) * 0.2
    h = 0.5 * (h + h.T) + numpy.diag(numpy.linspace(-0.6, 0.6, n))
    eri = fs.random_antisym(n, rng, hermitian=True) * 0.08
    en = numpy.diag(h).copy() + 0.03
    sysx = ftcc.GenericSystem(T, mu, h, eri, en)
    beta = 1.0 / T
    fo, fv = ftcc.ff(beta, en, mu), ftcc.ffv(beta, en, mu)
    thr = 0.1
    iocc = [i for i in range(n) if fo[i] > thr]
    ivir = [i for i in range(n) if fv[i] > thr]
    assert 0 < len(iocc) < n and 0 < len(ivir) < n and iocc != ivir
    fock, eri_t = sysx.g_fock_tot(), sysx.g_aint_tot()
    Fa, Ia = ftcc.ft_active_integrals(fock, eri_t, en, fo[iocc], fv[ivir],
                                      iocc, ivir)
    Ff, If = ftcc.ft_integrals(fock, eri_t, en, beta, mu)
    idx = {'o': iocc, 'v': ivir}
    for name in ('oo', 'ov', 'vo', 'vv'):
        ref = Ff[name][numpy.ix_(idx[name[0]], idx[name[1]])]
        assert numpy.array_equal(Fa[name], ref), name
    for name in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo', 'oovv', 'vooo',
                 'ooov', 'oooo'):
        ref = If[name][numpy.ix_(*(idx[x] for x in name))]
        assert numpy.abs(Ia[name] - ref).max() < 1e-15, name
    kw = dict(ngrid=6, damp=0.2, max_iter=200, econv=1e-12, tconv=1e-10)
    full = ftcc.ft_ccsd(sysx, T, mu, **kw)
    keep_all = ftcc.ft_ccsd(sysx, T, mu, athresh=1e-300, **kw)
    assert abs(full["Omega_cc"] - keep_all["Omega_cc"]) < 1e-14
    cut = ftcc.ft_ccsd(sysx, T, mu, athresh=thr, **kw)
    assert cut["T1"].shape[1:] == (len(ivir), len(iocc))
    assert abs(cut["Omega_cc"] - full["Omega_cc"]) > 1e-8


def test_ueg_params_and_mu():
    """The workload helpers: rs is the radius of the sphere that holds one
    electron (4/3 pi rs^3 N = L^3), E_F = k_F^2 / 2 with k_F^3 = 3 pi^2 N / L^3
    (the free-electron sphere holding N electrons of both spins), T = x E_F;
    solve_mu returns the mu at which the Fermi occupations sum to N."""
    N, rs, x = 14, 4.0, 0.05
    L, T, Ef = ftcc.ueg_params(19, rs, N, x)
    assert abs(4.0 / 3.0 * numpy.pi * rs ** 3 * N - L ** 3) < 1e-12 * L ** 3
    kF = (3.0 * numpy.pi ** 2 * N / L ** 3) ** (1.0 / 3.0)
    assert abs(Ef - 0.5 * kF * kF) < 1e-15 and abs(T - x * Ef) < 1e-17
    # kF from the electron count of the filled sphere: 2 (4/3 pi kF^3) /
    # (2 pi / L)^3 = N
    assert abs(2 * 4.0 / 3.0 * numpy.pi * kF ** 3 / (2 * numpy.pi / L) ** 3
               - N) < 1e-10
    s0 = ftcc.UEGSystem(T, L, 0.5 * (2 * numpy.pi / L) ** 2 * 8.0, 0.0,
                        norb=19)
    en = s0.g_energies_tot()
    mu = ftcc.solve_mu(en, 1.0 / T, N)
    assert abs(ftcc.ff(1.0 / T, en, mu).sum() - N) < 1e-10
    assert abs(ftcc.ff(1.0 / T, en, mu + 1e-3).sum() - N) > 1e-6


def test_fock_space_builders_closed_forms():
    """The brute-force checker's own pieces (oracle/fock_space.py), by what
    the mathematics fixes: the Jordan-Wigner matrices obey the canonical
    anticommutation relations; without interaction the grand potential is
    the free-fermion sum of GP0 over the eigenvalues of h, and with one
    pair interaction V = <01||01> the doubly occupied two-mode state has
    energy h00 + h11 + V; at T = 0 the projection gives the driving terms
    f[a,i] and <ab||ij>."""
    nso = 4
    ops = fs.jw_annihilators(nso)
    eye = numpy.eye(1 << nso)
    for p in range(nso):
        for q in range(nso):
            acr = ops[p] @ ops[q].T + ops[q].T @ ops[p]
            aa = ops[p] @ ops[q] + ops[q] @ ops[p]
            assert numpy.array_equal(acr, eye * (p == q))
            assert not aa.any()
    rng = numpy.random.default_rng(13)
    h = rng.normal(size=(nso, nso))
    h = 0.5 * (h + h.T)
    beta, mu = 1.7, 0.2
    eps = numpy.linalg.eigvalsh(h)
    om = fs.ft_fci_omega(h, numpy.zeros((nso,) * 4), beta, mu)
    assert abs(om - ftcc.GP0(beta, eps, mu).sum()) < 1e-12
    # two modes, one pair interaction
    h2 = numpy.diag([0.3, -0.2])
    V = 0.7
    eri = numpy.zeros((2, 2, 2, 2))
    eri[0, 1, 0, 1] = eri[1, 0, 1, 0] = V
    eri[0, 1, 1, 0] = eri[1, 0, 0, 1] = -V
    H = fs.hamiltonian(h2, eri, fs.jw_annihilators(2))
    assert abs(H[3, 3] - (0.3 - 0.2 + V)) < 1e-15
    assert abs(H[1, 1] - 0.3) < 1e-15 and abs(H[2, 2] + 0.2) < 1e-15
    # T = 0: <Phi_i^a|H|Phi_0> = f_ai, <Phi_ij^ab|H|Phi_0> = <ab||ij>
    no, nv = 2, 2
    n = no + nv
    f = rng.normal(size=(n, n))
    f = 0.5 * (f + f.T)
    g = fs.random_antisym(n, rng, hermitian=True)
    r1, r2 = fs.cc_projection(f, g, numpy.zeros((nv, no)),
                              numpy.zeros((nv, nv, no, no)), no)
    assert numpy.abs(r1 - f[no:, :no]).max() < 1e-12
    assert numpy.abs(r2 - g[no:, no:, :no, :no]).max() < 1e-12
