"""Parity of the HIP path (librbmpp.so through the C ABI) with the CPU oracle
on identical seeded inputs.  FP64 paths: relative 1e-10 (normwise where a
whole matrix is compared, scaled by the condition of the copy where noted);
sampler trajectories: bit-exact configurations."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

RTOL = 1e-10
GUARD = 1e4  # the engine's default conditioning guard


def _api():
    from rbm_pp_amd import api
    return api


def _model(L, alphar, alphac=0, J2=0.5, seed=1, sigma_w=0.05):
    spec = lt.j1j2_model(L, alphar=alphar, alphac=alphac, J2=J2)
    return spec, lt.random_params(spec, seed_w=seed, seed_f=seed + 100, sigma_w=sigma_w)


def _engine(api, spec, p, proj, **kw):
    """proj "block": the default normal projection (cos(pi/2) = 0, block-form
    kernels); "exact": the reference's weights bit for bit (skew-form
    kernels).  The oracle keeps the reference's weights in both cases."""
    eng = api.Engine(spec, params=p, **kw)
    if proj == "exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    return eng


PROJ = pytest.mark.parametrize("proj", ["block", "exact"])


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


CASES = [(4, 1, 0), (4, 2, 1), (6, 2, 0), (10, 4, 0)]


@PROJ
@pytest.mark.parametrize("L,ar,ac", CASES)
def test_eval_psi_pf_invm(oracle_lib, L, ar, ac, proj):
    api = _api()
    spec, p = _model(L, ar, ac)
    orc = oracle_lib.Oracle(spec, p)
    eng = _engine(api, spec, p, proj, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 6, seed=L)
    got = eng.eval(xs, with_invm=True)
    for s, x in enumerate(xs):
        st = orc.eval(x)
        assert _relerr(got["psi"][s], st["psi"]) < RTOL
        assert _relerr(got["pf"][s], st["pf"]) < RTOL
        for q in range(spec.C):
            inv_ref = st["invm"][q]
            cond = np.linalg.cond(inv_ref)
            assert _relerr(got["invm"][s, q], inv_ref) < max(RTOL, 1e-15 * cond), (s, q, cond)


@PROJ
@pytest.mark.parametrize("L,ar,ac", CASES)
def test_ratio_exchange(oracle_lib, L, ar, ac, proj):
    api = _api()
    spec, p = _model(L, ar, ac)
    orc = oracle_lib.Oracle(spec, p)
    eng = _engine(api, spec, p, proj, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 8, seed=100 + L)
    rng = np.random.default_rng(L)
    i1 = np.array([int(rng.integers(spec.N)) for _ in xs], np.int32)
    i2 = np.array([int(rng.choice(np.nonzero(x != x[a])[0])) for x, a in zip(xs, i1)], np.int32)
    pf, psi = eng.ratio_exchange(xs, i1, i2)
    for s, x in enumerate(xs):
        orc.eval(x)
        _, pfn, psin = orc.ratio_exchange(int(i1[s]), int(i2[s]))
        assert _relerr(pf[s], pfn) < RTOL
        assert abs(psi[s] - psin[0]) <= RTOL * abs(psin[0])


@PROJ
@pytest.mark.parametrize("L,ar,ac", CASES)
def test_local_energy_and_ok(oracle_lib, L, ar, ac, proj):
    api = _api()
    spec, p = _model(L, ar, ac)
    orc = oracle_lib.Oracle(spec, p)
    eng = _engine(api, spec, p, proj, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 6, seed=200 + L)
    got = eng.local(xs, want_o=True)
    for s, x in enumerate(xs):
        orc.eval(x)
        e = orc.eloc()
        o = orc.oloc()
        assert abs(got["eloc"][s] - e) <= RTOL * abs(e)
        assert _relerr(got["oloc"][s], o) < RTOL


@pytest.mark.parametrize("L,ar,ac,W,ns,K", [(4, 1, 0, 4, 6, 8), (6, 2, 1, 3, 4, 20), (10, 4, 0, 2, 3, 32),
                                             (6, 2, 0, 3, 6, 0), (10, 4, 0, 2, 4, 0)])
@PROJ
def test_sampler_trajectory_bit_exact(oracle_lib, L, ar, ac, W, ns, K, proj):
    """GPU walkers and the oracle chain driven by the same Philox streams
    visit the same configurations; amplitudes agree to FP64 rounding.  K = 0
    is the reference's adaptive recompute period (make_sample.f90:74-128),
    carried from call to call per walker."""
    api = _api()
    spec, p = _model(L, ar, ac, sigma_w=0.1)
    seed = 1234
    nwarm = 4
    eng = _engine(api, spec, p, proj, n_walkers=W, n_samples=ns, seed=seed, n_warmup=nwarm, recompute_k=K)
    eng.make_sample(True, shift=0.0)
    x_gpu, psi_gpu = eng.get_samples()
    wk = eng.get_walkers()
    orc = oracle_lib.Oracle(spec, p)
    states = []
    for w in range(W):
        rng = oracle_lib.Rng("philox", seed, stream=w)
        st = orc.new_state()
        out = orc.make_sample(rng, True, nsample=ns, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                              state=st, cond_guard=GUARD)
        states.append((rng, st, out["nfastup"]))
        assert np.array_equal(x_gpu[w * ns:(w + 1) * ns], out["x"]), f"walker {w}"
        # the two fast-update chains round differently; up to 32 accepts
        # between recomputes they agree to 1e-9, with the adaptive period (up
        # to 100 accepts) to 1e-8
        np.testing.assert_allclose(psi_gpu[w * ns:(w + 1) * ns], out["psi"], rtol=1e-9 if K else 1e-8)
        # after many fast updates: normwise over the copies of the walker
        assert _relerr(wk["pf"][w], out["pf"][-1]) < 1e-8
        assert wk["naccept"][w] == out["naccept"]
    # the X^-1 each walker holds after the launch (in block form an accepted
    # last proposal's update is applied at the end of the launch) against a
    # recompute of its final configuration, entry by entry.  The walker keeps
    # its own electron labels (Ele_Idx); rbmpp_eval numbers the electrons of
    # each spin in site order, so electron i of the walker is electron
    # rank[i] of the recompute and InvM_walker(i, j) = InvM_eval(rank[i], rank[j]).
    wi = eng.get_walkers(with_invm=True)
    ele = eng.walker_electrons()
    rec = eng.eval(wi["x"], with_invm=True)["invm"]
    H = spec.N // 2
    for w in range(W):
        e = ele[w]
        assert (wi["x"][w][e[:H]] == 1).all() and (wi["x"][w][e[H:]] == -1).all()
        rank = np.empty(spec.N, np.int64)
        rank[:H] = np.argsort(np.argsort(e[:H]))
        rank[H:] = H + np.argsort(np.argsort(e[H:]))
        for q in range(rec.shape[1]):
            want = rec[w, q][np.ix_(rank, rank)]
            assert _relerr(wi["invm"][w, q], want) < 1e-8, f"walker {w} copy {q} X^-1"
    # a second call continues every chain from its last sample
    eng.make_sample(False, shift=1e-3)
    x2, _ = eng.get_samples()
    for w in range(W):
        rng, st, nf = states[w]
        out = orc.make_sample(rng, False, nsample=ns, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                              shift=1e-3, state=st, cond_guard=GUARD, nfastup=nf)
        assert np.array_equal(x2[w * ns:(w + 1) * ns], out["x"]), f"walker {w} (second call)"


@pytest.mark.parametrize("L,ar,ac", [(4, 2, 1), (6, 2, 0), (10, 4, 0)])
@PROJ
def test_measure_s_and_sr_step(oracle_lib, L, ar, ac, proj):
    api = _api()
    spec, p = _model(L, ar, ac)
    orc = oracle_lib.Oracle(spec, p)
    n = 96
    xs = _configs(spec.N, n, seed=300 + L)
    eng = _engine(api, spec, p, proj, n_walkers=1, n_samples=n)
    eng.set_samples(xs)
    shift = 2e-3
    got = eng.calc_phys_quantity_opt(shift=shift)
    ref = orc.measure(xs, shift=shift, want_s=True)
    norm = ref["norm"]
    assert got["norm"] == pytest.approx(norm, rel=RTOL)
    assert got["Etot"] == pytest.approx(ref["Etot"] / norm, rel=RTOL)
    assert got["Etot2"] == pytest.approx(ref["Etot2"] / norm, rel=RTOL)
    assert got["wf_av"] == pytest.approx(ref["wf_av"] / n, rel=RTOL)
    assert _relerr(got["gvec"], ref["gvec"] / norm) < RTOL
    assert _relerr(got["ovec"], ref["ovec"] / norm) < RTOL
    ob = ref["ovec"] / norm
    S_ref = ref["smat"] / norm - np.outer(ob, ob)
    S = eng.smat()
    assert _relerr(S, S_ref) < 1e-9
    gam0 = orc.gamma()
    rc, gam_ref, rep_ref, _ = orc.sr_step(ref["Etot"] / norm, norm, ref["gvec"], ref["ovec"], ref["smat"],
                                          gam0, s_factor0=1e-3)
    assert rc == 0
    gam, rep = eng.sr_step(gam0, s_factor0=1e-3, refine_iters=1)
    assert rep["residual"] < 1e-12
    assert rep["lambda_"] == pytest.approx(rep_ref[0], rel=1e-8)
    assert _relerr(gam - gam0, gam_ref - gam0) < 1e-8
    np.testing.assert_allclose(eng.gamma(), gam, rtol=0, atol=0)


@pytest.mark.parametrize("max_change,clipped", [(1e-4, True), (0.04, False)])
def test_sr_step_clip_branch_10x10(oracle_lib, max_change, clipped):
    """Both branches of the max-change clip (SR_pdposv.f90:329-354) on the
    10x10 workload at delta_tau = 0.05 (delW_max = 1.0e-3 here, so
    max_change = 1e-4 clips): the HIP step, delW_max and lambda against the
    oracle.
    The two solve the same shifted system in different orders, so the bound
    is the system's condition number times a few units of rounding:
    tol = max(1e-10, 50 eps cond(A)), with A the shifted matrix of the
    oracle's own S (stated in the assertion message)."""
    api = _api()
    spec, p = _model(10, 4, 0)
    orc = oracle_lib.Oracle(spec, p)
    n = 96
    xs = _configs(spec.N, n, seed=910)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=n)
    eng.set_samples(xs)
    shift = 2e-3
    eng.calc_phys_quantity_opt(shift=shift)
    ref = orc.measure(xs, shift=shift, want_s=True)
    norm = ref["norm"]
    gam0 = orc.gamma()
    rc, gam_ref, rep_ref, x_ref = orc.sr_step(ref["Etot"] / norm, norm, ref["gvec"], ref["ovec"], ref["smat"],
                                              gam0, s_factor0=1e-3, delta_tau=0.05, max_change=max_change)
    assert rc == 0
    assert (rep_ref[1] > max_change) == clipped
    gam, rep = eng.sr_step(gam0, s_factor0=1e-3, delta_tau=0.05, max_change=max_change, refine_iters=1)
    # condition of the scaled, shifted system (SR_pdposv.f90:18-170)
    ob = ref["ovec"] / norm
    S = ref["smat"] / norm - np.outer(ob, ob)
    nvr = spec.nv_rbm
    D = np.ones(spec.nv); D[nvr:] = rep_ref[2]
    A = S * np.outer(D, D)
    A[np.diag_indices(spec.nv)] += rep_ref[3] * rep_ref[4]
    tol = max(1e-10, 50 * np.finfo(float).eps * np.linalg.cond(A))
    assert rep["delw_max"] == pytest.approx(rep_ref[1], rel=tol), tol
    assert rep["lambda_"] == pytest.approx(rep_ref[0], rel=tol), tol
    assert _relerr(gam - gam0, gam_ref - gam0) < tol, tol
    if clipped:
        dg = (gam - gam0)[:nvr]
        assert np.abs(dg).max() <= max_change * (1 + 1e-12)


@pytest.mark.parametrize("L,ar,ac", [(4, 2, 1), (6, 2, 0), (10, 4, 0)])
def test_fused_measurement_matches_recompute(oracle_lib, L, ar, ac):
    """Measurement fused into the sampler (fast-updated X^-1) against the
    oracle's calc_phys_quantity_opt (X^-1 recomputed per sample) on the same
    samples: equal to the fast-update drift over up to 16 accepts (bounds
    1e-8 for norm and energy, 5e-8 normwise for g, <O> and S, whose
    O entries carry X^-1 itself)."""
    api = _api()
    spec, p = _model(L, ar, ac, sigma_w=0.1)
    W, ns = 8, 6
    eng = api.Engine(spec, n_walkers=W, n_samples=ns, seed=77, n_warmup=5, recompute_k=16, params=p,
                     fused_measure=True)
    shift = 1e-3
    eng.make_sample(True, shift=shift)
    got = eng.calc_phys_quantity_opt(shift=shift)
    xs, _ = eng.get_samples()
    orc = oracle_lib.Oracle(spec, p)
    ref = orc.measure(xs, shift=shift, want_s=True)
    norm = ref["norm"]
    assert got["norm"] == pytest.approx(norm, rel=1e-8)
    assert got["Etot"] == pytest.approx(ref["Etot"] / norm, rel=1e-8)
    assert _relerr(got["gvec"], ref["gvec"] / norm) < 5e-8
    assert _relerr(got["ovec"], ref["ovec"] / norm) < 5e-8
    ob = ref["ovec"] / norm
    S_ref = ref["smat"] / norm - np.outer(ob, ob)
    S = eng.smat()
    assert _relerr(S, S_ref) < 5e-8
    # the f_ii rows of S are exactly zero, as the SR step's checks require
    # (SR_pdposv.f90:18-40): the fast-updated X^-1 keeps a zero diagonal
    d = np.diag(S)
    assert (d[:spec.nv_rbm] > 0).all()
    assert (d[spec.nv_rbm:] > 0).sum() == spec.n_slater - spec.n_slater // spec.N
    # a different shift falls back to the recompute path (chain-drawn samples:
    # the energy sum cancels partly, bound 1e-9)
    got2 = eng.calc_phys_quantity_opt(shift=2e-3)
    ref2 = orc.measure(xs, shift=2e-3, want_s=False)
    assert got2["Etot"] == pytest.approx(ref2["Etot"] / ref2["norm"], rel=1e-9)
    eng.make_sample(False, shift=shift)
    eng.calc_phys_quantity_opt(shift=shift)
    eng.sr_step(eng.gamma(), s_factor0=1e-3)


@pytest.mark.parametrize("prec,bound", [("S_FP32", 1e-5), ("S_BF16X3", 5e-5)])
def test_s_build_mixed_error_bound(oracle_lib, prec, bound):
    """Mixed S on the matrix cores: centred in FP64, products in FP32 or in
    split BF16 x 3 (hi*hi + hi*lo + lo*hi), FP32 accumulation; stated bounds
    normwise relative to the FP64 S on this workload (nv = 218: a partial
    128-column tile)."""
    api = _api()
    spec, p = _model(6, 2, 0)
    n = 256
    xs = _configs(spec.N, n, seed=7)
    e64 = api.Engine(spec, n_walkers=1, n_samples=n, params=p, s_precision=api.S_FP64)
    emx = api.Engine(spec, n_walkers=1, n_samples=n, params=p, s_precision=getattr(api, prec))
    for e in (e64, emx):
        e.set_samples(xs)
        e.calc_phys_quantity_opt(shift=1e-3)
    S64, Smx = e64.smat(), emx.smat()
    err = np.linalg.norm(Smx - S64) / np.linalg.norm(S64)
    assert err < bound, err
    assert np.array_equal(Smx, Smx.T)


def test_vmc_energy_4x4_against_enumeration(oracle_lib):
    """GPU sampler + estimator: VMC energy of a fixed 4x4 wave function within
    5 error bars of its exact expectation (12870-state enumeration)."""
    import ed
    api = _api()
    spec, p = _model(4, 1, 0, sigma_w=0.1)
    orc = oracle_lib.Oracle(spec, p)
    states, Hg = ed.hamiltonian(4, 1.0, 0.5, gauge=True)
    m = orc.measure(states.astype(np.int32), shift=0.0, want_s=False)
    psi = m["psi"]
    e_exact = (np.conj(psi) @ (Hg @ psi)).real / (np.abs(psi) ** 2).sum()
    W, ns = 64, 100
    eng = api.Engine(spec, n_walkers=W, n_samples=ns, seed=99, n_warmup=20, recompute_k=32, params=p)
    eng.make_sample(True)
    r = eng.calc_phys_quantity_opt(shift=0.0)
    loc = eng.local(eng.get_samples()[0], want_o=False)
    el = (loc["eloc"] / loc["psi"]).real
    err = el.reshape(W, ns).mean(1).std() / np.sqrt(W)
    assert abs(r["Etot"] - e_exact) < 5 * err + 1e-12


def test_guarded_chain_through_near_singular_configuration(oracle_lib):
    """Walker 354 of this 10x10 workload passes close to a singular
    configuration of one projection copy in its second call (see the oracle
    test of the same chain).  With the conditioning guard the GPU chain
    visits the same configurations as the guarded oracle chain and its
    amplitudes stay exact to 1e-6."""
    api = _api()
    spec = lt.j1j2_model(10, alphar=4, J2=0.5)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    W, ns, w = 360, 49, 354
    eng = api.Engine(spec, n_walkers=W, n_samples=ns, seed=1, n_warmup=30, recompute_k=32,
                     params=p, cond_guard=GUARD)
    eng.make_sample(True)
    eng.make_sample(False)
    xs, psi = eng.get_samples()
    orc = oracle_lib.Oracle(spec, p)
    rng = oracle_lib.Rng("philox", 1, stream=w)
    st = orc.new_state()
    orc.make_sample(rng, True, nsample=ns, nwarmup=30, fixed_k=32, ref_draws=False, state=st,
                    cond_guard=GUARD)
    out = orc.make_sample(rng, False, nsample=ns, nwarmup=30, fixed_k=32, ref_draws=False,
                          state=st, cond_guard=GUARD)
    assert np.array_equal(xs[w * ns:(w + 1) * ns], out["x"])
    exact = eng.local(xs[w * ns:(w + 1) * ns], want_o=False)["psi"]
    assert _relerr(psi[w * ns:(w + 1) * ns] / exact, np.ones(ns)) < 1e-6


@pytest.mark.parametrize("L,ar,ac", [(4, 1, 0), (4, 1, 1)])
def test_autograd_logpsi_against_finite_differences(L, ar, ac):
    """d ln|psi| / d gamma from the hand-written adjoint (torch backward of
    rbm_pp_amd.autograd) against central differences of the GPU forward,
    h = 1e-6: relative 1e-6 of the largest component."""
    import torch
    api = _api()
    from rbm_pp_amd.autograd import log_abs_psi
    spec, p = _model(L, ar, ac, sigma_w=0.1)
    eng = api.Engine(spec, n_walkers=1, n_samples=1, params=p)
    xs = _configs(spec.N, 5, seed=42)
    w = torch.tensor(np.linspace(0.5, 1.5, len(xs)))
    gam = torch.tensor(p.gamma(), requires_grad=True)
    y = (w * log_abs_psi(gam, eng, xs)).sum()
    y.backward()
    grad = gam.grad.numpy()
    h = 1e-6
    g0 = p.gamma()
    fd = np.zeros_like(g0)
    for k in range(len(g0)):
        for sgn in (1, -1):
            gk = g0.copy()
            gk[k] += sgn * h
            fd[k] += sgn * float((w * log_abs_psi(torch.tensor(gk), eng, xs)).sum()) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=0, atol=1e-6 * np.abs(fd).max())
