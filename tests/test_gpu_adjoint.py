"""The adjoint of the FT-CCSD fixed point on the GPU: the transposed
residual and integrator against the oracle, the lambda amplitudes and
E, S, N against kelvin's printed lambda-equation values, finite differences
of full solves and the oracle's adjoint, and the autograd wrapper."""
import json
import os

import numpy
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from oracle import ftcc as oracle  # noqa: E402
from oracle import fock_space as fs  # noqa: E402

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_goldens.json")))


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def to_np(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return numpy.abs(a - b).max() / max(numpy.abs(b).max(), 1e-300)


def random_problem(no, nv, rng):
    n = no + nv
    f = rng.normal(size=(n, n)) * 0.3
    eri = fs.random_antisym(n, rng, hermitian=False) * 0.1
    sl = {'o': slice(0, no), 'v': slice(no, n)}
    F = oracle.Blocks(oo=f[sl['o'], sl['o']], ov=f[sl['o'], sl['v']],
                      vo=f[sl['v'], sl['o']], vv=f[sl['v'], sl['v']])
    I = oracle.Blocks({k: numpy.ascontiguousarray(
        eri[sl[k[0]], sl[k[1]], sl[k[2]], sl[k[3]]])
        for k in ('vvvv', 'vvvo', 'vovv', 'vvoo', 'vovo', 'oovv', 'vooo',
                  'ooov', 'oooo')})
    return F, I


def antisym(x):
    x = x - numpy.swapaxes(x, -4, -3)
    return x - numpy.swapaxes(x, -2, -1)


def _engine(F, I, no, nv, ng, rng):
    from kelvin_amd.engine import FTEngine
    from kelvin_amd.ov_blocks import one_e_blocks, two_e_blocks
    eng = FTEngine(no, nv, ng)
    eng.set_blocks(one_e_blocks(**{k: F[k] for k in ('oo', 'ov', 'vo', 'vv')}),
                   two_e_blocks(**{k: I[k] for k in I}))
    eng.set_denominators(rng.normal(size=(nv, no)),
                         rng.normal(size=(nv, nv, no, no)))
    eng.set_grid(numpy.zeros(ng), numpy.zeros((ng, ng)))
    return eng


@pytest.mark.parametrize("no,nv,ng", [(3, 4, 2), (5, 5, 3), (9, 11, 2)])
def test_transposed_residual_matches_oracle(no, nv, ng):
    """ftcc_adjoint_apply: J^T of the residual at every grid point (the
    kernel of the lambda equations) against the oracle's exact transpose,
    projected onto antisymmetric doubles, on random non-Hermitian blocks
    with o != v; with every grid point's intermediates kept and with none
    kept (rebuilt from the amplitudes at each application)."""
    d = dev()
    rng = numpy.random.default_rng(31 * no + nv)
    F, I = random_problem(no, nv, rng)
    t1 = rng.normal(size=(ng, nv, no)) * 0.2
    t2 = antisym(rng.normal(size=(ng, nv, nv, no, no))) * 0.1
    v1 = rng.normal(size=(ng, nv, no))
    v2 = antisym(rng.normal(size=(ng, nv, nv, no, no)))
    ref1 = numpy.empty_like(v1)
    ref2 = numpy.empty_like(v2)
    for y in range(ng):
        g1, g2 = oracle.stanton_rhs_transpose(F, I, t1[y], t2[y], v1[y],
                                              v2[y])
        ref1[y], ref2[y] = g1, oracle.antisymmetrize(g2)
    T1, T2 = (torch.as_tensor(x, device=d) for x in (t1, t2))
    V1, V2 = (torch.as_tensor(x, device=d) for x in (v1, v2))
    for cache_gb in ("64", "0"):
        os.environ["FTCC_ADJOINT_CACHE_GB"] = cache_gb
        try:
            eng = _engine(F, I, no, nv, ng, rng)
            eng.adjoint_begin(T1, T2)
            assert eng.adjoint_cached() == (ng if cache_gb == "64" else 0)
            G1, G2 = eng.empty_amplitudes()
            eng.adjoint_apply(V1, V2, G1, G2)
            eng.adjoint_end()
        finally:
            del os.environ["FTCC_ADJOINT_CACHE_GB"]
        assert rel(to_np(G1), ref1) < 1e-12
        assert rel(to_np(G2), ref2) < 1e-12
        eng.close()


@pytest.mark.parametrize("quad", ['lin', 'sin', 'mid', 'L'])
def test_adjoint_integration_matches_oracle(quad):
    """ftcc_integrate_adjoint is the oracle's transpose of the integrator
    (uniform grids by the running power, others by one exp per pair), with
    denominators from orbital energies."""
    from kelvin_amd import quadrature
    from kelvin_amd.engine import FTEngine
    d = dev()
    rng = numpy.random.default_rng(41)
    ng, n = 9, 6
    ti, g, G = quadrature.ft_quad(ng, 4.0, quad)
    en = numpy.sort(rng.uniform(-1, 1, n))
    eng = FTEngine(n, n, ng)
    eri = fs.random_antisym(n, rng, hermitian=True) * 0.1
    eng.set_system(en, numpy.diag(en), eri, 1.0 / 0.25, 0.0)
    eng.set_grid(ti, G)
    l1 = rng.normal(size=(ng, n, n))
    l2 = rng.normal(size=(ng, n, n, n, n))
    D1, D2 = oracle.D1(en, en), oracle.D2(en, en)
    L1, L2 = (torch.as_tensor(x, device=d) for x in (l1, l2))
    O1, O2 = eng.empty_amplitudes()
    eng.integrate_adjoint(L1, L2, O1, O2)
    assert rel(to_np(O1), oracle.int_adjoint(ng, l1, ti, D1, G)) < 1e-13
    assert rel(to_np(O2), oracle.int_adjoint(ng, l2, ti, D2, G)) < 1e-13
    eng.close()


def _ueg_gold(key):
    from kelvin_amd.ueg_system import UEGSystem
    s = GOLD[key]["system"]
    L = 2 * numpy.pi if s["L"] == "2*pi" else s["L"]
    return UEGSystem(s["T"], L, s["Emax"], mu=s["mu"], norb=s["norb"],
                     orbtype='g'), s


@pytest.mark.parametrize("key,precision", [("G4", "fp64"), ("G4", "bf16x3"),
                                           ("G3", "fp64"), ("G5", "fp64")])
def test_esn_goldens(key, precision):
    """compute_ESN against the E, S, N that kelvin's lambda-equation
    compute_ESN printed (examples/ueg_ft_cc_compare.out:46-48,
    examples/ueg_ft_ccsd_ESN.out:43-45, bench/ueg_ft_ccsd_ESN19/
    ulambda_19_04_17.out:42-44), at 1e-6."""
    from kelvin_amd.ccsd import ccsd
    dev()
    g = GOLD[key]
    sysg, s = _ueg_gold(key)
    cc = ccsd(sysg, T=s["T"], mu=s["mu"], precision=precision, **g["ccsd"])
    cc.run()
    E, S, N = cc.compute_ESN()
    print("ESN %s %s: dE %.2e dS %.2e dN %.2e lambda %d it, %.2f s"
          % (key, precision, E - g["E"], S - g["S"], N - g["N"],
             cc.lambda_iter, cc.time_esn))
    assert cc.lambda_converged
    assert abs(E - g["E"]) < 1e-6
    assert abs(S - g["S"]) < 1e-6
    assert abs(N - g["N"]) < 1e-6
    assert abs(cc.N0 + cc.N1 + cc.Ncc - N) < 1e-12
    assert abs(cc.E0 + cc.E1 + cc.Ecc - E) < 1e-12


def test_refined_low_precision_lambda(monkeypatch):
    """The refined lambda iteration of a bf16x3 solve with every unbatched
    block of J^T as a split-bf16 product (threshold 0): one application
    differs from FP64 at the split-bf16 level, and the three phases end at
    the FP64 fixed point -- E, S, N within 1e-9 of the FP64 lambda solve
    from the same T."""
    from kelvin_amd.ccsd import ccsd
    from kelvin_amd.adjoint import solve_lambda
    dev()
    monkeypatch.setenv("FTCC_ADJOINT_LOWP_MIN_GFLOP", "0")
    g = GOLD["G4"]
    sysg, s = _ueg_gold("G4")
    kw = dict(g["ccsd"])
    kw["tconv"] = 1e-10
    cc = ccsd(sysg, T=s["T"], mu=s["mu"], precision="bf16x3", **kw)
    cc.run()
    eng = cc.engine
    E, S, N = cc.compute_ESN()
    assert cc.lambda_phases[1] > 0 and cc.lambda_phases[3] > 0
    assert cc.lambda_converged
    L1, L2 = cc.L1.clone(), cc.L2.clone()
    # the same lambda solve in FP64 throughout
    from kelvin_amd import adjoint as adj
    orig = adj.solve_lambda

    def fp64_only(*a, **k):
        k["low"] = None
        return orig(*a, **k)
    monkeypatch.setattr(adj, "solve_lambda", fp64_only)
    E64, S64, N64 = cc.compute_ESN()
    print("refined lambda: phases %s, dE %.2e dS %.2e dN %.2e"
          % (cc.lambda_phases, E - E64, S - S64, N - N64))
    assert abs(E - E64) < 1e-9
    assert abs(S - S64) < 1e-9
    assert abs(N - N64) < 1e-9
    # one application at the converged lambda: low against FP64
    T1, T2 = cc.T1, cc.T2
    eng.adjoint_begin(T1, T2)
    try:
        Lt1, Lt2 = eng.empty_amplitudes()
        eng.integrate_adjoint(L1, L2, Lt1, Lt2)
        out = []
        for p in ("fp64", "bf16x3"):
            eng.adjoint_precision(p)
            G1, G2 = eng.empty_amplitudes()
            eng.adjoint_apply(Lt1, Lt2, G1, G2)
            out.append((to_np(G1), to_np(G2)))
        eng.adjoint_precision("fp64")
    finally:
        eng.adjoint_end()
    d2 = rel(out[1][1], out[0][1])
    print("one application, bf16x3 against FP64: rel %.2e" % d2)
    assert 1e-12 < d2 < 1e-4
    assert rel(out[1][0], out[0][0]) < 1e-4


def test_esn_and_lambda_match_oracle_G4():
    """The G4 system: E, S, N and the lambda amplitudes in kelvin's layout
    equal the oracle's adjoint (kelvin's lambda iteration restated), the
    lambda iteration taking the same number of steps."""
    from kelvin_amd.ccsd import ccsd
    dev()
    g = GOLD["G4"]
    sysg, s = _ueg_gold("G4")
    cc = ccsd(sysg, T=s["T"], mu=s["mu"], **g["ccsd"])
    cc.run()
    E, S, N = cc.compute_ESN()
    osys = oracle.UEGSystem(s["T"], 2 * numpy.pi, s["Emax"], s["mu"],
                            norb=s["norb"])
    r = oracle.ft_ccsd_esn(osys, s["T"], s["mu"], **g["ccsd"])
    assert cc.lambda_iter == r["lambda_iter"]
    assert abs(E - r["E"]) < 1e-10
    assert abs(S - r["S"]) < 1e-9
    assert abs(N - r["N"]) < 1e-10
    assert rel(to_np(cc.L1), r["L1"]) < 1e-8
    assert rel(to_np(cc.L2), r["L2"]) < 1e-8
    # warm start from the converged lambda amplitudes: one iteration
    cc.compute_ESN(L1=cc.L1.clone(), L2=cc.L2.clone())
    assert cc.lambda_iter == 1
    assert abs(cc.E - E) < 1e-9


@pytest.mark.parametrize("kind", ["df", "generic", "pueg"])
def test_esn_matches_finite_differences(kind):
    """The adjoint against central differences of full solves (h = 1e-4)
    on the density-fitted model, on a system whose Fock matrix depends
    on the occupations (kelvin's g_fock_d_den term) and on the
    spin-polarised gas (kelvin/tests/test_ft_deriv.py:387-441), and against
    the oracle's adjoint (1e-10); gderiv=False against the oracle too."""
    from kelvin_amd.ccsd import ccsd
    from kelvin_amd.model_systems import ArraySystem, DFModelSystem
    from kelvin_amd.ueg_system import PUEGSystem
    dev()
    rng = numpy.random.default_rng(5)
    if kind == "df":
        T, mu = 0.3, 0.05
        sysg = DFModelSystem(T, mu, norb=3, seed=1, sigma=0.2)
        osys = oracle.DFModelSystem(T, mu, norb=3, seed=1, sigma=0.2)
    elif kind == "pueg":
        T, mu = 0.1, 0.1
        sysg = PUEGSystem(T, 2 * numpy.pi, 1.2, mu=mu, norb=7)
        osys = oracle.PUEGSystem(T, 2 * numpy.pi, 1.2, mu, norb=7)
    else:
        n = 6
        h = rng.normal(size=(n, n)) * 0.2
        h = 0.5 * (h + h.T) + numpy.diag(numpy.linspace(-0.6, 0.6, n))
        eri = fs.random_antisym(n, rng, hermitian=True) * 0.08
        en = numpy.diag(h).copy() + 0.03
        T, mu = 0.25, -0.02
        sysg = ArraySystem(T, mu, h, eri, en)
        osys = oracle.GenericSystem(T, mu, h, eri, en)
    kw = dict(ngrid=8, damp=0.2, max_iter=200, econv=1e-13, tconv=1e-11)
    cc = ccsd(sysg, T=T, mu=mu, **kw)
    cc.run()
    E, S, N = cc.compute_ESN()
    r = oracle.ft_ccsd_esn(osys, T, mu, **kw)
    assert abs(E - r["E"]) < 1e-10
    assert abs(S - r["S"]) < 1e-9
    assert abs(N - r["N"]) < 1e-10
    Ea, Sa, Na = cc.compute_ESN(gderiv=False)
    ra = oracle.ft_ccsd_esn(osys, T, mu, gderiv=False, **kw)
    assert abs(Ea - ra["E"]) < 1e-10
    assert abs(Na - ra["N"]) < 1e-10
    fd = ccsd(sysg, T=T, mu=mu, **kw)
    fd.run()
    Ef, Sf, Nf = fd.compute_ESN_fd(delta=1e-4, fd_econv=1e-13)
    assert abs(E - Ef) < 1e-7
    assert abs(S - Sf) < 1e-7
    assert abs(N - Nf) < 1e-7


def test_autograd_grand_potential():
    """kelvin_amd.adjoint.grand_potential: Omega(beta, mu) with the adjoint
    as its backward; the gradient equals (E - mu N - Omega)/beta and -N of
    compute_ESN and central differences of the forward (h = 1e-4)."""
    from kelvin_amd.adjoint import grand_potential
    from kelvin_amd.model_systems import DFModelSystem
    dev()

    def make(T, mu):
        return DFModelSystem(T, mu, norb=3, seed=1, sigma=0.2)
    kw = dict(ngrid=8, damp=0.2, max_iter=200, econv=1e-13, tconv=1e-11)
    beta = torch.tensor(1.0 / 0.3, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    om = grand_potential(make, beta, mu, **kw)
    om.backward()
    h = 1e-4

    def f(b, m):
        return float(grand_potential(make, torch.tensor(b, dtype=torch.float64),
                                     torch.tensor(m, dtype=torch.float64), **kw))
    b0, m0 = float(beta), float(mu)
    gb = (f(b0 + h, m0) - f(b0 - h, m0)) / (2 * h)
    gm = (f(b0, m0 + h) - f(b0, m0 - h)) / (2 * h)
    assert abs(float(beta.grad) - gb) < 1e-7 * max(1.0, abs(gb))
    assert abs(float(mu.grad) - gm) < 1e-7


def test_transposed_residual_spin_blocks_equal_dense():
    """On a spin-orbital UEG (19 k-points) the transposed residual runs
    over the spin blocks its operands' rules allow; it equals the dense
    evaluation (FTCC_SPIN=0) to rounding, for amplitudes after a step and
    the integrated lambda guess."""
    from kelvin_amd import quadrature
    from kelvin_amd.engine import FTEngine
    from kelvin_amd.ueg_system import ueg_from_config
    dev()
    sysg = ueg_from_config(19, 2.0, N=14, t_over_tf=0.3)
    en = sysg.g_energies_tot()
    n, ng = en.shape[0], 4
    beta = 1.0 / sysg.T
    ti, g, G = quadrature.ft_quad(ng, beta, 'lin')
    out = []
    for spin in ("1", "0"):
        os.environ["FTCC_SPIN"] = spin
        try:
            eng = FTEngine(n, n, ng)
            fock, eri, _ = sysg.g_device_integrals(eng)
            eng.set_system(en, fock, eri, beta, sysg.mu)
        finally:
            del os.environ["FTCC_SPIN"]
        del fock, eri
        eng.set_grid(ti, G)
        assert (eng.spin_mode() != 0) == (spin == "1")
        T1, T2 = eng.empty_amplitudes()
        eng.driving_terms(T1, T2)
        eng.integrate(T1, T2)
        B1, B2 = eng.empty_amplitudes()
        eng.step(T1, T2, B1, B2, g, beta, 0.3)
        L1, L2 = eng.empty_amplitudes()
        eng.lambda_guess(T1, g, beta, L1, L2)
        Lt1, Lt2 = eng.empty_amplitudes()
        eng.integrate_adjoint(L1, L2, Lt1, Lt2)
        eng.adjoint_begin(T1, T2)
        G1, G2 = eng.empty_amplitudes()
        eng.adjoint_apply(Lt1, Lt2, G1, G2)
        eng.adjoint_end()
        out.append((to_np(G1), to_np(G2)))
        eng.close()
    assert rel(out[0][0], out[1][0]) < 1e-12
    assert rel(out[0][1], out[1][1]) < 1e-12


@pytest.mark.parametrize("spin", [False, True])
def test_density1_matches_oracle(spin):
    """ftcc_density1 (kelvin's ft_cc_equations.ccsd_1rdm on the device)
    against the oracle's restatement on random amplitudes and lambda
    amplitudes, with orbital-energy denominators on a Simpson grid; dense,
    and spin-blocked on a spin-conserving system (amplitudes projected onto
    the allowed blocks)."""
    from kelvin_amd import quadrature
    from kelvin_amd.engine import FTEngine
    d = dev()
    rng = numpy.random.default_rng(43)
    ng, n, beta = 7, 6, 3.0
    ti, g, G = quadrature.ft_quad(ng, beta, 'lin')
    en = numpy.sort(rng.uniform(-1, 1, n))
    eng = FTEngine(n, n, ng)
    eri = fs.random_antisym(n, rng, hermitian=True) * 0.1
    lab = numpy.arange(n) % 2
    if spin:
        # zero every spin-forbidden integral so that blocking applies
        ok = ((lab[:, None, None, None] + lab[None, :, None, None])
              == (lab[None, None, :, None] + lab[None, None, None, :]))
        eri = eri * ok
    eng.set_system(en, numpy.diag(en), eri, beta, 0.0)
    if spin:
        eng.set_spin(lab, lab)
    eng.set_grid(ti, G)
    t1 = rng.normal(size=(ng, n, n)) * 0.3
    t2 = antisym(rng.normal(size=(ng, n, n, n, n))) * 0.2
    l1 = rng.normal(size=(ng, n, n)) * 0.3
    l2 = antisym(rng.normal(size=(ng, n, n, n, n))) * 0.2
    if spin:
        m1 = lab[:, None] == lab[None, :]
        m2 = ((lab[:, None, None, None] + lab[None, :, None, None])
              == (lab[None, None, :, None] + lab[None, None, None, :]))
        t1, l1 = t1 * m1, l1 * m1
        t2, l2 = t2 * m2, l2 * m2
    D1, D2 = oracle.D1(en, en), oracle.D2(en, en)
    ref = oracle.ccsd_1rdm(t1, t2, l1, l2, D1, D2, ti, ng, g, G)
    T1, T2 = (torch.as_tensor(x, device=d) for x in (t1, t2))
    Lam1, Lam2 = eng.lambda_from_kelvin(l1, l2, g, beta)
    Lt1, Lt2 = eng.empty_amplitudes()
    eng.integrate_adjoint(Lam1, Lam2, Lt1, Lt2)
    got = eng.density1(T1, T2, Lt1, Lt2, g, beta)
    for name, a, b in zip(("pia", "pba", "pji", "pai"), got, ref):
        print("density1 %s %s: rel %.2e" % (spin, name, rel(to_np(a), b)))
        assert rel(to_np(a), b) < 1e-12
    eng.close()


def test_full_1rdm_matches_oracle_G4():
    """ccsd.full_1rdm on the G4 system: the unrelaxed density of the GPU
    lambda solve plus the Fermi occupations, against the oracle's (its own
    converged T and L, kelvin/ccsd.py:1333-1387 and 1961-2006)."""
    from kelvin_amd.ccsd import ccsd
    dev()
    g = GOLD["G4"]
    sysg, s = _ueg_gold("G4")
    cc = ccsd(sysg, T=s["T"], mu=s["mu"], **g["ccsd"])
    cc.run()
    rdm = cc.full_1rdm()
    osys = oracle.UEGSystem(s["T"], 2 * numpy.pi, s["Emax"], s["mu"],
                            norb=s["norb"])
    r = oracle.ft_ccsd_esn(osys, s["T"], s["mu"], **g["ccsd"])
    beta = 1.0 / s["T"]
    en = osys.g_energies_tot()
    ng = g["ccsd"]["ngrid"]
    ti, gw, G = oracle.ft_quad(ng, beta, 'lin')
    D1, D2 = oracle.D1(en, en), oracle.D2(en, en)
    blocks = oracle.ccsd_1rdm(r["T1"], r["T2"], r["L1"], r["L2"], D1, D2, ti,
                              ng, gw, G)
    fo, fv = oracle.ff(beta, en, s["mu"]), oracle.ffv(beta, en, s["mu"])
    ref = oracle.normal_1rdm(*blocks, fo, fv, beta)
    # the correlation part alone (the occupations on the diagonal are the
    # same numbers on both sides)
    cor, cref = rdm - numpy.diag(fo), ref - numpy.diag(fo)
    print("full_1rdm G4: n=%d, |correlation part| %.3e, rel %.2e, blocks "
          "%s, trace %.10f" % (len(en), numpy.abs(cref).max(), rel(cor, cref),
                               ["%.2e" % numpy.abs(b).max() for b in blocks],
                               numpy.trace(rdm)))
    assert numpy.abs(cref).max() > 1e-6
    assert rel(cor, cref) < 1e-7


def test_full_1rdm_unrestricted_blocks():
    """full_1rdm on the G4 system built with 'u' orbitals: the alpha and
    beta blocks of the 'g' system's density (kelvin/ccsd.py:1961-1988 returns
    [alpha, beta] for 'u'); the off-diagonal spin blocks of the 'g' density
    vanish (the density conserves spin)."""
    from kelvin_amd.ccsd import ccsd
    from kelvin_amd.ueg_system import UEGSystem
    dev()
    g = GOLD["G4"]
    sysg, s = _ueg_gold("G4")
    L = 2 * numpy.pi if s["L"] == "2*pi" else s["L"]
    sysu = UEGSystem(s["T"], L, s["Emax"], mu=s["mu"], norb=s["norb"],
                     orbtype='u')
    ccg = ccsd(sysg, T=s["T"], mu=s["mu"], **g["ccsd"])
    ccg.run()
    rg = ccg.full_1rdm()
    ccu = ccsd(sysu, T=s["T"], mu=s["mu"], **g["ccsd"])
    ccu.run()
    ra, rb = ccu.full_1rdm()
    na = ra.shape[0]
    assert rb.shape == ra.shape and 2 * na == rg.shape[0]
    assert numpy.abs(ra - rg[:na, :na]).max() < 1e-10
    assert numpy.abs(rb - rg[na:, na:]).max() < 1e-10
    assert numpy.abs(rg[:na, na:]).max() < 1e-12
    assert numpy.abs(rg[na:, :na]).max() < 1e-12
    # the two spins of the unpolarised gas carry the same density
    assert numpy.abs(ra - rb).max() < 1e-10
