"""The launch-shape tuner (rbmpp_tune): it times the S build's output tiles
and split-K counts and the per-sample estimator's workgroup counts on the
handle's own data and keeps the fastest.  Its choices change no per-sample
value and no measured sum; S changes only by the order in which the split
partials are added, and the SR step by that through the solve."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu


def _relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("prec_name,s_tol", [("fp64", 1e-13), ("bf16x3", 1e-5)])
def test_tuned_launch_shapes_keep_the_results(prec_name, s_tol):
    from rbm_pp_amd import api
    prec = {"fp64": api.S_FP64, "bf16x3": api.S_BF16X3}[prec_name]
    spec = lt.j1j2_model(6, alphar=2, J2=0.5)
    p = lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.05)
    W, ns = 512, 8
    eng = api.Engine(spec, params=p, n_walkers=W, n_samples=ns, seed=11, n_warmup=4, s_precision=prec)
    eng.make_sample(True, shift=0.0)
    r0 = eng.calc_phys_quantity_opt(shift=1e-3)
    S0 = eng.smat()
    g0, _ = eng.sr_step(p.gamma(), s_factor0=1e-3)
    eng.set_gamma(p.gamma())  # the step set the new parameters; measure the stored samples at the old ones

    rep = eng.tune(api.TUNE_S_BUILD | api.TUNE_MEASURE)
    # S-build candidates: 128-wide tiles (and 256-wide for split BF16 x 3),
    # split-K counts clamped to the rows
    assert rep["s_tile"] in ((128, 256) if prec_name == "bf16x3" else (128,))
    assert rep["s_nsplit"] >= 1 and rep["s_build_s"] > 0
    assert rep["s_default_tile"] in (128, 256) and rep["s_default_nsplit"] >= 1
    assert rep["measure_grid"] in (256, 512, 1024, 2048, 4096) and rep["measure_s"] > 0
    assert rep["measure_configurations"] == W * ns

    # the estimator grid only assigns configurations to workgroups: the
    # measured sums are those of the untuned launch
    r1 = eng.calc_phys_quantity_opt(shift=1e-3)
    for k in ("Etot", "Etot2", "norm", "wf_av"):
        assert r1[k] == pytest.approx(r0[k], rel=1e-14, abs=0.0), k
    assert _relerr(r1["gvec"], r0["gvec"]) < 1e-14
    assert _relerr(r1["ovec"], r0["ovec"]) < 1e-14
    S1 = eng.smat()
    assert _relerr(S1, S0) < s_tol
    assert np.array_equal(S1, S1.T)
    g1, rep1 = eng.sr_step(p.gamma(), s_factor0=1e-3)
    assert rep1["residual"] < 1e-12
    assert _relerr(g1 - p.gamma(), g0 - p.gamma()) < 1e-9
