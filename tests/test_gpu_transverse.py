"""The transverse_code variant on the GPU: spin-projection weights of
transverse_code/spin_projection.f90:30-38 selected through
rbmpp_set_projection, against the oracle with the same weights (itself pinned
to the transverse reference routines by tests/test_oracle.py).  With these
weights both projection copies have full same-spin blocks (cos*sin = -1, +1),
where the normal weights leave them zero or at rounding level, so the
Pfaffian, ratio, update and estimator paths all see a different matrix
structure.  Bounds as in test_gpu_parity.py."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

RTOL = 1e-10
GUARD = 1e4


def _api():
    from rbm_pp_amd import api
    return api


def _setup(oracle_lib, L, ar, ac, sigma_w=0.05, **kw):
    api = _api()
    spec = lt.j1j2_model(L, alphar=ar, alphac=ac, J2=0.5)
    p = lt.random_params(spec, seed_w=5, seed_f=105, sigma_w=sigma_w)
    orc = oracle_lib.Oracle(spec, p)
    orc.set_projection(1)
    eng = api.Engine(spec, params=p, **kw)
    eng.set_projection(api.PROJ_TRANSVERSE)
    return api, spec, p, orc, eng


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


@pytest.mark.parametrize("L,ar,ac", [(4, 2, 1), (6, 2, 0), (10, 4, 0)])
def test_eval_ratio_local_energy_and_ok(oracle_lib, L, ar, ac):
    api, spec, p, orc, eng = _setup(oracle_lib, L, ar, ac, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 6, seed=40 + L)
    got = eng.eval(xs, with_invm=True)
    loc = eng.local(xs, want_o=True)
    rng = np.random.default_rng(L)
    i1 = np.array([int(rng.integers(spec.N)) for _ in xs], np.int32)
    i2 = np.array([int(rng.choice(np.nonzero(x != x[a])[0])) for x, a in zip(xs, i1)], np.int32)
    pf2, psi2 = eng.ratio_exchange(xs, i1, i2)
    for s, x in enumerate(xs):
        st = orc.eval(x)
        assert _relerr(got["psi"][s], st["psi"]) < RTOL
        assert _relerr(got["pf"][s], st["pf"]) < RTOL
        for q in range(spec.C):
            inv_ref = st["invm"][q]
            cond = np.linalg.cond(inv_ref)
            assert _relerr(got["invm"][s, q], inv_ref) < max(RTOL, 1e-15 * cond), (s, q, cond)
        e = orc.eloc()
        assert abs(loc["eloc"][s] - e) <= RTOL * abs(e)
        assert _relerr(loc["oloc"][s], orc.oloc()) < RTOL
        _, pfn, psin = orc.ratio_exchange(int(i1[s]), int(i2[s]))
        assert _relerr(pf2[s], pfn) < RTOL
        assert abs(psi2[s] - psin[0]) <= RTOL * abs(psin[0])


def test_switching_back_restores_the_normal_weights(oracle_lib):
    api, spec, p, orc, eng = _setup(oracle_lib, 4, 2, 1, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 3, seed=9)
    tr = eng.eval(xs)["psi"]
    eng.set_projection(api.PROJ_NORMAL)
    nm = eng.eval(xs)["psi"]
    orc.set_projection(0)
    for s, x in enumerate(xs):
        assert _relerr(nm[s], orc.eval(x)["psi"]) < RTOL
    assert _relerr(tr, nm) > 1e-3  # the two variants are different wave functions
    with pytest.raises(RuntimeError):
        eng.set_projection(3)


@pytest.mark.parametrize("L,ar,ac,W,ns,K", [(4, 1, 0, 4, 6, 8), (6, 2, 1, 3, 4, 20), (10, 4, 0, 2, 3, 32)])
def test_sampler_trajectory_bit_exact(oracle_lib, L, ar, ac, W, ns, K):
    """Same Philox streams: the same configurations, bit for bit."""
    seed, nwarm = 4321, 4
    api, spec, p, orc, eng = _setup(oracle_lib, L, ar, ac, sigma_w=0.1, n_walkers=W, n_samples=ns, seed=seed,
                                    n_warmup=nwarm, recompute_k=K)
    eng.make_sample(True, shift=0.0)
    x_gpu, psi_gpu = eng.get_samples()
    wk = eng.get_walkers()
    for w in range(W):
        rng = oracle_lib.Rng("philox", seed, stream=w)
        st = orc.new_state()
        out = orc.make_sample(rng, True, nsample=ns, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                              state=st, cond_guard=GUARD)
        assert np.array_equal(x_gpu[w * ns:(w + 1) * ns], out["x"]), f"walker {w}"
        # fast-update drift: the oracle's own chain is 5.7e-9 off a recompute
        # of the same samples on the 6x6 case; both chains against the
        # recomputed amplitudes, bound 2e-8
        psi_re = np.array([np.ravel(orc.eval(x)["psi"])[0] for x in out["x"]])
        np.testing.assert_allclose(psi_gpu[w * ns:(w + 1) * ns], psi_re, rtol=2e-8)
        np.testing.assert_allclose(out["psi"], psi_re, rtol=2e-8)
        assert _relerr(wk["pf"][w], out["pf"][-1]) < 1e-8
        assert wk["naccept"][w] == out["naccept"]


@pytest.mark.parametrize("L,ar,ac", [(4, 2, 1), (10, 4, 0)])
def test_fused_measurement_and_sr_step(oracle_lib, L, ar, ac):
    api, spec, p, orc, eng = _setup(oracle_lib, L, ar, ac, sigma_w=0.1, n_walkers=8, n_samples=6, seed=78,
                                    n_warmup=5, recompute_k=16, fused_measure=True)
    shift = 3e-3
    eng.make_sample(True, shift=shift)
    got = eng.calc_phys_quantity_opt(shift=shift)
    xs, _ = eng.get_samples()
    ref = orc.measure(xs, shift=shift, want_s=True)
    norm = ref["norm"]
    assert got["norm"] == pytest.approx(norm, rel=1e-8)
    assert got["Etot"] == pytest.approx(ref["Etot"] / norm, rel=1e-8)
    assert _relerr(got["gvec"], ref["gvec"] / norm) < 1e-8
    ob = ref["ovec"] / norm
    S_ref = ref["smat"] / norm - np.outer(ob, ob)
    assert _relerr(eng.smat(), S_ref) < 1e-8
    gam0 = orc.gamma()
    rc, gam_ref, rep_ref, _ = orc.sr_step(ref["Etot"] / norm, norm, ref["gvec"], ref["ovec"], ref["smat"],
                                          gam0, s_factor0=1e-3)
    assert rc == 0
    gam, rep = eng.sr_step(gam0, s_factor0=1e-3, refine_iters=1)
    assert _relerr(gam - gam0, gam_ref - gam0) < 1e-6


def test_vmc_energy_4x4_against_enumeration(oracle_lib):
    """Sampler + estimator with the transverse weights: VMC energy within 5
    error bars of the exact expectation of the same wave function."""
    import ed
    api, spec, p, orc, eng = _setup(oracle_lib, 4, 1, 0, sigma_w=0.1, n_walkers=64, n_samples=100, seed=98,
                                    n_warmup=20, recompute_k=32)
    states, Hg = ed.hamiltonian(4, 1.0, 0.5, gauge=True)
    m = orc.measure(states.astype(np.int32), shift=0.0, want_s=False)
    psi = m["psi"]
    e_exact = (np.conj(psi) @ (Hg @ psi)).real / (np.abs(psi) ** 2).sum()
    eng.make_sample(True)
    r = eng.calc_phys_quantity_opt(shift=0.0)
    loc = eng.local(eng.get_samples()[0], want_o=False)
    el = (loc["eloc"] / loc["psi"]).real
    err = el.reshape(64, 100).mean(1).std() / np.sqrt(64)
    assert abs(r["Etot"] - e_exact) < 5 * err + 1e-12
