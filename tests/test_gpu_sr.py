"""The SR step's device pipeline (scalars, D S D + shift, Cholesky, blocked
substitutions, refinement, lambda) replayed from its captured graph against
eager launches; measurements loaded through rbmpp_set_measured against
rbmpp_measure; the SR step at the 16x16 workload's size (Nv = 5136) on a
synthetic measurement in the three S precisions."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu


def _configs(N, n, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((n, N), np.int32)
    for k in range(n):
        x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
        rng.shuffle(x)
        out[k] = x
    return out


def _relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def test_sr_step_graph_replays_match_eager():
    """Each measurement is solved eagerly and by a graph replay; the two are
    bit-identical (same kernels, same order, no atomics).  Eager work and
    synchronisations (another measurement, local energies) run between the
    replays."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(10, alphar=4, J2=0.5)
    p = lt.random_params(spec, seed_w=1, seed_f=101, sigma_w=0.05)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=96)
    gam0 = p.gamma()
    for rep in range(4):
        xs = _configs(spec.N, 96, seed=700 + rep)
        eng.set_gamma(gam0)
        eng.set_samples(xs)
        eng.calc_phys_quantity_opt(shift=1e-3)
        for refine in (1, 2) if rep == 3 else (1,):
            ge, re_ = eng.sr_step(gam0, s_factor0=1e-3 * (rep + 1), refine_iters=refine, graph=False)
            gg, rg = eng.sr_step(gam0, s_factor0=1e-3 * (rep + 1), refine_iters=refine, graph=True)
            np.testing.assert_array_equal(gg, ge)
            assert rg == re_
            assert rg["residual"] < 1e-12
        eng.local(xs[:8], want_o=True)


@pytest.mark.parametrize("L,ar,ac", [(10, 4, 0), (4, 2, 1)])
def test_set_measured_reproduces_measure(L, ar, ac):
    """The per-configuration values of a measurement (psi, reweighted E_loc,
    rw and O rows rw O_loc) loaded through rbmpp_set_measured give the same
    S and SR step as rbmpp_measure on the same samples."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(L, alphar=ar, alphac=ac, J2=0.5)
    p = lt.random_params(spec, seed_w=2, seed_f=3, sigma_w=0.05)
    n = 128
    xs = _configs(spec.N, n, seed=31)
    shift = 2e-3
    a = api.Engine(spec, params=p, n_walkers=1, n_samples=n)
    a.set_samples(xs)
    a.calc_phys_quantity_opt(shift=shift)
    psi, el, rw = a.get_measured()
    S_a = a.smat()
    g_a, rep_a = a.sr_step(p.gamma(), s_factor0=1e-3)
    a.set_gamma(p.gamma())  # the step loaded new parameters
    loc = a.local(xs, want_o=True)
    O = rw[:, None] * loc["oloc"]
    if ac == 0:
        rows = O.real
    else:
        rows = np.empty((2 * n, spec.nv))
        rows[0::2], rows[1::2] = O.real, O.imag
    b = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    b.set_measured(n, psi=psi, eloc=el, rw=rw, orows=rows)
    assert _relerr(b.smat(), S_a) < 1e-13
    g_b, rep_b = b.sr_step(p.gamma(), s_factor0=1e-3)
    assert _relerr(g_b - p.gamma(), g_a - p.gamma()) < 1e-9
    assert rep_b["lambda_"] == pytest.approx(rep_a["lambda_"], rel=1e-10)


def test_sr_step_16x16_synthetic_three_precisions():
    """The SR step at the 16x16 workload's parameter count (alphar = 16,
    2x2-sublattice pair product: Nv = 5136) on a synthetic measurement of
    8192 configurations: every precision solves the FP64 system (residual
    < 1e-10; the mixed precisions by conjugate gradients on the operator of
    the O rows), S is exactly symmetric, the mixed-precision S stays within
    stated bounds of FP64 (FP32 1e-5, BF16 x 3 1e-4 normwise) and the mixed
    steps within 1e-8 of the FP64 step."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(16, alphar=16, J2=0.55)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    assert spec.nv == 5136
    n = 8192
    eng.set_measured(n, seed=5)
    res = {}
    for name, prec in (("fp64", api.S_FP64), ("fp32", api.S_FP32), ("bf16x3", api.S_BF16X3)):
        eng.set_s_precision(prec)
        S = eng.smat()
        g, rep = eng.sr_step(p.gamma(), s_factor0=1e-4, refine_iters=1 if name == "fp64" else 3)
        assert rep["residual"] < 1e-10, (name, rep["residual"])
        assert np.array_equal(S, S.T)
        res[name] = (S, g - p.gamma())
    S64, d64 = res["fp64"]
    for name, bound in (("fp32", 1e-5), ("bf16x3", 1e-4)):
        err = np.linalg.norm(res[name][0] - S64) / np.linalg.norm(S64)
        assert err < bound, (name, err)
        assert np.linalg.norm(res[name][1] - d64) / np.linalg.norm(d64) < 1e-8, name


@pytest.mark.parametrize("prec", ["S_FP32", "S_BF16X3"])
def test_mixed_step_solved_on_fp64_operator_10x10(prec):
    """On a real 10x10 measurement (1024 walkers x 8 samples), the SR step
    with S from FP32 or split-BF16 x 3 products, solved by conjugate
    gradients on the FP64 operator of the O rows (preconditioned by the
    assembled matrix's factor; more iterations until the residual is below
    1e-11), equals the FP64 step to 1e-8 and its lambda to 1e-10, whether
    zero or two iterations are captured; the same step refined against the
    assembled matrix stays at the S error's distance (about 1e-3 and 4e-3)."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(10, alphar=4, J2=0.5)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    eng = api.Engine(spec, params=p, n_walkers=1024, n_samples=8, seed=3, s_precision=api.S_FP64)
    eng.make_sample(True, shift=0.0)
    eng.calc_phys_quantity_opt(shift=0.0)
    gam = p.gamma()
    g64, r64 = eng.sr_step(gam, s_factor0=1e-4)
    d64 = g64 - gam
    assert r64["cg_iters"] == 0
    eng.set_s_precision(getattr(api, prec))
    for captured in (0, 2):
        gm, rm = eng.sr_step(gam, s_factor0=1e-4, refine_iters=captured)
        err = np.linalg.norm(gm - g64) / np.linalg.norm(d64)
        print(f"{prec}: {captured} captured, {rm['cg_iters']} iterations, residual {rm['residual']:.1e}, "
              f"step error {err:.2e}, boost {rm['precond_boost']}", flush=True)
        assert rm["residual"] < 1e-11
        assert err < 1e-8
        assert rm["lambda_"] == pytest.approx(r64["lambda_"], rel=1e-10)
    ga, ra = eng.sr_step(gam, s_factor0=1e-4, refine_iters=2, assembled_refine=True)
    err_a = np.linalg.norm(ga - g64) / np.linalg.norm(d64)
    print(f"{prec}: refined against the assembled S: step error {err_a:.2e}", flush=True)
    assert ra["cg_iters"] == 0
    assert err_a > 100 * err


@pytest.mark.parametrize("L,ar,ac,proj", [(10, 4, 0, "block"), (10, 4, 0, "exact"), (6, 2, 1, "block"),
                                          (12, 2, 0, "block")])
def test_measurement_run_to_run_identical(L, ar, ac, proj):
    """Two measurements of the same samples give bit-identical O rows, S and
    SR step: the pair-product O_k are summed per orbital by one owning thread
    in a fixed order (no atomics), every reduction has a fixed order.  Block
    form, the skew form with the reference's exact weights, a complex model,
    and N > 104 (the global-scratch estimator)."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(L, alphar=ar, alphac=ac, J2=0.5)
    p = lt.random_params(spec, seed_w=4, seed_f=5, sigma_w=0.05)
    n = 64
    xs = _configs(spec.N, n, seed=900 + L)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=n)
    if proj == "exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    out = []
    for rep in range(2):
        eng.set_gamma(p.gamma())
        eng.set_samples(xs)
        r = eng.calc_phys_quantity_opt(shift=1e-3)
        S = eng.smat()
        g, _ = eng.sr_step(p.gamma(), s_factor0=1e-3)
        loc = eng.local(xs[:16], want_o=True)
        out.append((r, S, g, loc))
    (r0, S0, g0, l0), (r1, S1, g1, l1) = out
    assert r0["Etot"] == r1["Etot"] and r0["norm"] == r1["norm"]
    for k in ("psi", "eloc", "oloc"):
        np.testing.assert_array_equal(l0[k], l1[k])
    np.testing.assert_array_equal(S0, S1)
    np.testing.assert_array_equal(g0, g1)
