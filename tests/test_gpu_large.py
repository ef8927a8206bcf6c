"""Parity of the N > 104 paths with the CPU oracle: lattices 12x12, 14x14
and the 16x16 workload (J2/J1 = 0.55, alphar = 16, C4v + parity, N = 256).
Above N = 104 the normal projection's sampler keeps F^-1 of each copy with
two columns per lane and inverts F in LDS (k_sweep<256, true>); the
reference's exact weights keep X^-1 in folded skew storage with the recompute
matrix in global scratch (k_sweep<256, false>, wg_skew_inverse).  The
normal projection's estimator is the block-form k_meas_blk_big (F of each
copy inverted in LDS by the workgroup); the exact weights' is the skew-form
k_local (rbmpp_local.hip).  Bounds as
in test_gpu_parity: FP64 values relative 1e-10 (normwise per matrix, scaled
by the copy's condition where noted), trajectories bit-exact."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

RTOL = 1e-10
GUARD = 1e4

# (L, alphar, J2): 12x12 and 14x14 at the uneven-batch shape, 16x16 at the
# large workload's
LARGE = [(12, 2, 0.5), (14, 2, 0.4), (16, 16, 0.55)]


def _api():
    from rbm_pp_amd import api
    return api


def _model(L, alphar, J2, seed=1, sigma_w=0.05):
    spec = lt.j1j2_model(L, alphar=alphar, J2=J2)
    return spec, lt.random_params(spec, seed_w=seed, seed_f=seed + 100, sigma_w=sigma_w)


def _configs(N, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
        rng.shuffle(x)
        out.append(x)
    return np.array(out)


def _relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("L,ar,J2", LARGE)
def test_eval_large(oracle_lib, L, ar, J2):
    api = _api()
    spec, p = _model(L, ar, J2)
    orc = oracle_lib.Oracle(spec, p)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 3, seed=L)
    got = eng.eval(xs, with_invm=True)
    for s, x in enumerate(xs):
        st = orc.eval(x)
        assert _relerr(got["psi"][s], st["psi"]) < RTOL
        assert _relerr(got["pf"][s], st["pf"]) < RTOL
        for q in range(spec.C):
            inv_ref = st["invm"][q]
            cond = np.linalg.cond(inv_ref)
            assert _relerr(got["invm"][s, q], inv_ref) < max(RTOL, 1e-15 * cond), (s, q, cond)


@pytest.mark.parametrize("L,ar,J2", LARGE)
def test_ratio_exchange_large(oracle_lib, L, ar, J2):
    api = _api()
    spec, p = _model(L, ar, J2)
    orc = oracle_lib.Oracle(spec, p)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 4, seed=100 + L)
    rng = np.random.default_rng(L)
    i1 = np.array([int(rng.integers(spec.N)) for _ in xs], np.int32)
    i2 = np.array([int(rng.choice(np.nonzero(x != x[a])[0])) for x, a in zip(xs, i1)], np.int32)
    pf, psi = eng.ratio_exchange(xs, i1, i2)
    for s, x in enumerate(xs):
        orc.eval(x)
        _, pfn, psin = orc.ratio_exchange(int(i1[s]), int(i2[s]))
        assert _relerr(pf[s], pfn) < RTOL
        # psi(x') is the sum of the 16 copies' terms, which cancel in part:
        # relative to the sum of their magnitudes
        assert abs(psi[s] - psin[0]) <= RTOL * np.abs(psin[1:]).sum()


@pytest.mark.parametrize("L,ar,J2,proj", [(L, ar, J2, "block") for L, ar, J2 in LARGE] + [(12, 2, 0.5, "exact")])
def test_local_energy_and_ok_large(oracle_lib, L, ar, J2, proj):
    """proj "block": the block-form estimator (k_meas_blk_big); "exact": the
    reference's weights, the skew-form estimator (k_local)."""
    api = _api()
    spec, p = _model(L, ar, J2)
    orc = oracle_lib.Oracle(spec, p)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    if proj == "exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    xs = _configs(spec.N, 3, seed=200 + L)
    got = eng.local(xs, want_o=True)
    for s, x in enumerate(xs):
        orc.eval(x)
        e = orc.eloc()
        o = orc.oloc()
        assert abs(got["eloc"][s] - e) <= RTOL * abs(e)
        assert _relerr(got["oloc"][s], o) < RTOL


@pytest.mark.parametrize("L,ar,J2,K,proj", [(12, 2, 0.5, 32, "block"), (14, 2, 0.4, 0, "block"),
                                            (16, 16, 0.55, 32, "block"), (16, 16, 0.55, 0, "block"),
                                            (12, 2, 0.5, 32, "exact"), (16, 16, 0.55, 0, "exact")])
def test_sampler_trajectory_large(oracle_lib, L, ar, J2, K, proj):
    """Two walkers, two calls, against the oracle chain on the same Philox
    streams; configurations bit-exact, X^-1 after the launch entrywise
    against a recompute in the walker's electron order.  proj "block": the
    default normal projection (F^-1 per copy, two columns per lane, F
    inverted in LDS); "exact": the reference's weights (folded skew X^-1)."""
    api = _api()
    spec, p = _model(L, ar, J2, sigma_w=0.1)
    W, ns, nwarm, seed = 2, 2, 2, 4321
    eng = api.Engine(spec, params=p, n_walkers=W, n_samples=ns, seed=seed, n_warmup=nwarm, recompute_k=K)
    if proj == "exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    eng.make_sample(True, shift=0.0)
    x_gpu, psi_gpu = eng.get_samples()
    wk = eng.get_walkers(with_invm=True)
    ele = eng.walker_electrons()
    orc = oracle_lib.Oracle(spec, p)
    states, drift = [], []
    for w in range(W):
        rng = oracle_lib.Rng("philox", seed, stream=w)
        st = orc.new_state()
        out = orc.make_sample(rng, True, nsample=ns, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                              state=st, cond_guard=GUARD)
        states.append((rng, st, out["nfastup"]))
        assert np.array_equal(x_gpu[w * ns:(w + 1) * ns], out["x"]), f"walker {w}"
        np.testing.assert_allclose(psi_gpu[w * ns:(w + 1) * ns], out["psi"], rtol=1e-8)
        assert _relerr(wk["pf"][w], out["pf"][-1]) < 1e-8
        assert wk["naccept"][w] == out["naccept"]
        # the oracle chain's own fast-update drift of X^-1 at this point (the
        # next call re-evaluates from x, so the recompute changes nothing)
        drift.append(orc.recompute(st)[1])
    rec = eng.eval(wk["x"], with_invm=True)["invm"]
    H = spec.N // 2
    for w in range(W):
        e = ele[w]
        rank = np.empty(spec.N, np.int64)
        rank[:H] = np.argsort(np.argsort(e[:H]))
        rank[H:] = H + np.argsort(np.argsort(e[H:]))
        # the fast-updated X^-1 drifts from the recompute by the rounding of
        # up to K (adaptive: 100) rank-2 updates, growing with N and the
        # conditioning of the copies; the GPU chain may drift as much as the
        # oracle's chain of the same moves, within a factor 10 (the
        # reference's own drift threshold is 1e-5, make_sample.f90:94-113)
        bound = max(1e-8, 10 * drift[w])
        for q in range(spec.C):
            want = rec[w, q][np.ix_(rank, rank)]
            assert _relerr(wk["invm"][w, q], want) < bound, (f"walker {w} copy {q} X^-1", drift[w])
    eng.make_sample(False, shift=1e-3)
    x2, _ = eng.get_samples()
    for w in range(W):
        rng, st, nf = states[w]
        out = orc.make_sample(rng, False, nsample=ns, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                              shift=1e-3, state=st, cond_guard=GUARD, nfastup=nf)
        assert np.array_equal(x2[w * ns:(w + 1) * ns], out["x"]), f"walker {w} (second call)"


def _sr_solution_numpy(spec, m, s_factor0):
    """SR_pdposv.f90:4-323 with numpy's dense solve (the oracle's own step is
    pinned against the same rebuild in tests/test_oracle.py)."""
    nv, nvr, norm = spec.nv, spec.nv_rbm, m["norm"]
    ovec = m["ovec"] / norm
    etot = m["Etot"] / norm
    S = m["smat"] / norm - np.outer(ovec, ovec)
    g = m["gvec"] / norm - 2 * etot * ovec
    d = np.diag(S)
    scal = np.sqrt(d[:nvr][d[:nvr] > 0].mean() / d[nvr:][d[nvr:] > 0].mean())
    D = np.ones(nv); D[nvr:] = scal
    A = S * np.outer(D, D)
    Sav = np.diag(A)[np.diag(A) > 0].mean()
    A[np.diag_indices(nv)] += Sav * max(1e-4, s_factor0)
    xs_ = np.linalg.solve(A, g * D)
    lam = np.sqrt(xs_ @ A @ xs_) * 5 / spec.L
    if lam > 1:
        xs_ = xs_ / lam
    ev = np.linalg.eigvalsh(A)  # A is symmetric positive definite
    return xs_ * D, lam, ev[-1] / ev[0]


@pytest.mark.parametrize("L,ar,J2", [(12, 2, 0.5), (16, 16, 0.55)])
def test_measure_s_and_sr_step_large(oracle_lib, L, ar, J2):
    api = _api()
    spec, p = _model(L, ar, J2)
    orc = oracle_lib.Oracle(spec, p)
    n = 32
    xs = _configs(spec.N, n, seed=300 + L)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=n)
    eng.set_samples(xs)
    shift = 2e-3
    got = eng.calc_phys_quantity_opt(shift=shift)
    ref = orc.measure(xs, shift=shift, want_s=True)
    norm = ref["norm"]
    assert got["norm"] == pytest.approx(norm, rel=RTOL)
    assert got["Etot"] == pytest.approx(ref["Etot"] / norm, rel=RTOL)
    assert got["Etot2"] == pytest.approx(ref["Etot2"] / norm, rel=RTOL)
    assert _relerr(got["gvec"], ref["gvec"] / norm) < RTOL
    assert _relerr(got["ovec"], ref["ovec"] / norm) < RTOL
    ob = ref["ovec"] / norm
    S_ref = ref["smat"] / norm - np.outer(ob, ob)
    assert _relerr(eng.smat(), S_ref) < 1e-9
    gam0 = orc.gamma()
    x_np, lam, cond = _sr_solution_numpy(spec, ref, 1e-3)
    tol = max(1e-10, 50 * np.finfo(float).eps * cond)
    gam, rep = eng.sr_step(gam0, s_factor0=1e-3, refine_iters=1)
    assert rep["residual"] < 1e-12
    assert rep["lambda_"] == pytest.approx(lam, rel=tol), tol
    # gamma rebuilt from the numpy solution (SR_pdposv.f90:329-354)
    step = 0.05 * x_np
    dmax = np.abs(step).max()
    if dmax > 0.04:
        step = step * (0.04 / dmax)
    want = gam0 - step
    f = want[spec.nv_rbm:]
    want[spec.nv_rbm:] = f * (0.25 / np.abs(f).mean())
    assert _relerr(gam - gam0, want - gam0) < tol, tol


def test_adjoint_ok_6x6_256_configurations(oracle_lib):
    """The 6x6 workload (alphar = 2, C4v + parity): O_k from the hand-written
    adjoint on 256 seeded configurations against the oracle's O_k (relative
    1e-10) and against central finite differences of the GPU amplitude,
    h = 1e-6 (d psi / d gamma_k; relative 2e-8 of each configuration's
    largest component: the difference's rounding error is about 5e-10 of it
    on this workload)."""
    api = _api()
    spec, p = _model(6, 2, 0.5, sigma_w=0.1)
    orc = oracle_lib.Oracle(spec, p)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 256, seed=2561)
    got = eng.local(xs, want_o=True)
    O = got["oloc"]
    for s, x in enumerate(xs):
        orc.eval(x)
        assert _relerr(O[s], orc.oloc()) < RTOL, s
    g0 = p.gamma()
    h = 1e-6
    fd = np.zeros((len(xs), spec.nv), np.complex128)
    for k in range(spec.nv):
        gp = g0.copy(); gp[k] += h
        gm = g0.copy(); gm[k] -= h
        eng.set_gamma(gp)
        a = eng.eval(xs)["psi"][:, 0]
        eng.set_gamma(gm)
        b = eng.eval(xs)["psi"][:, 0]
        fd[:, k] = (a - b) / (2 * h)
    eng.set_gamma(g0)
    scale = np.abs(O).max(axis=1, keepdims=True)
    err = (np.abs(O - fd) / scale).max()
    assert err < 2e-8, err
