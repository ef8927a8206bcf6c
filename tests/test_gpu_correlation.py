"""The correlation run on the GPU (rbmpp_measure_correlation: Chi_sp, Chi_vbs,
wf_sgn of Calc_Correlation/calc_phys_quantity.f90:264-550, normalised as
Calc_Correlation/main.f90:543-581) against the oracle on the same samples.
The GPU forms the pair ratios in block form (default normal projection) from
Yh = Gdd F^-1, Xh = F^-1 Guu and W = Yh Guu, in skew form (the reference's
exact weights, transverse weights) from Y = G InvM^T and Z = G Y^T; the
oracle (pinned to the reference's calc_local_correlation and to exact
enumeration) from the rank-2 update of pfupdate_two.f90 per pair, so the bound
is relative 1e-9 normwise, like the S comparisons."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu


def _api():
    from rbm_pp_amd import api
    return api


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


def _check(got, ref, n):
    norm = ref["norm"]
    assert got["norm"] == pytest.approx(norm, rel=1e-10)
    assert got["Etot"] == pytest.approx(ref["Etot"] / norm, rel=1e-10)
    assert got["Etot2"] == pytest.approx(ref["Etot2"] / norm, rel=1e-10)
    assert got["wf_av"] == pytest.approx(ref["wf_av"] / n, rel=1e-10)
    assert abs(got["wf_sgn"] - ref["wf_sgn"] / n) <= 1e-12
    assert _relerr(got["chi_sp"][:, 0], ref["chi_sp"][:, 0] / norm) < 1e-10
    assert _relerr(got["chi_sp"][:, 1], ref["chi_sp"][:, 1] / norm) < 1e-9
    assert _relerr(got["chi_vbs"], ref["chi_vbs"] / norm) < 1e-9


@pytest.mark.parametrize("L,ar,ac,n", [(4, 2, 1, 24), (6, 2, 0, 16), (6, 1, 1, 12), (10, 4, 0, 6)])
@pytest.mark.parametrize("proj", ["block", "exact"])
def test_correlation_matches_oracle(oracle_lib, L, ar, ac, n, proj):
    """proj "block": the default normal projection (cos(pi/2) = 0, block-form
    kernel); "exact": the reference's weights bit for bit (skew-form kernel).
    The oracle keeps the reference's weights in both cases."""
    api = _api()
    spec = lt.j1j2_model(L, alphar=ar, alphac=ac, J2=0.5)
    p = lt.random_params(spec, seed_w=L, seed_f=L + 100, sigma_w=0.05)
    xs = _configs(spec.N, n, seed=500 + L)
    eng = api.Engine(spec, n_walkers=1, n_samples=n, params=p)
    if proj == "exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    eng.set_samples(xs)
    shift = 1e-3
    got = eng.measure_correlation(shift=shift)
    ref = oracle_lib.Oracle(spec, p).measure_correlation(xs, shift=shift)
    _check(got, ref, n)
    # the energy part equals the plain energy measurement
    e = eng.calc_phys_quantity_after_opt(shift=shift)
    assert got["Etot"] == pytest.approx(e["Etot"], rel=1e-13)


def test_correlation_transverse_weights(oracle_lib):
    api = _api()
    spec = lt.j1j2_model(6, alphar=2, alphac=1, J2=0.5)
    p = lt.random_params(spec, seed_w=9, seed_f=109, sigma_w=0.05)
    xs = _configs(spec.N, 10, seed=9)
    eng = api.Engine(spec, n_walkers=1, n_samples=10, params=p)
    eng.set_projection(api.PROJ_TRANSVERSE)
    eng.set_samples(xs)
    orc = oracle_lib.Oracle(spec, p)
    orc.set_projection(1)
    _check(eng.measure_correlation(shift=0.0), orc.measure_correlation(xs, shift=0.0), 10)


def test_correlation_closed_forms_10x10_sampled():
    """On sampler-drawn 10x10 samples (2048 of them): Chi_sp(.,ii) = 1 and
    sum_{i<j} Chi_sp(1,.) = -N/2 (total Sz = 0) exactly up to rounding, the
    bond-summed Chi_vbs diagonal is positive, and SzSz is symmetric under
    the lattice translations on average (the wave function is projected on
    them), to the statistical error."""
    api = _api()
    spec = lt.j1j2_model(10, alphar=4, alphac=0, J2=0.5)
    p = lt.random_params(spec, seed_w=1, seed_f=101, sigma_w=0.05)
    W, ns = 256, 8
    eng = api.Engine(spec, n_walkers=W, n_samples=ns, seed=5, n_warmup=20, recompute_k=32, params=p)
    eng.make_sample(True, shift=0.0)
    r = eng.measure_correlation(shift=1e-5)
    N = spec.N
    diag = np.array([i * (i + 1) // 2 + i for i in range(N)])
    np.testing.assert_allclose(r["chi_sp"][diag], 1.0, rtol=1e-12)
    off = np.ones(N * (N + 1) // 2, bool)
    off[diag] = False
    assert r["chi_sp"][off, 0].sum() == pytest.approx(-N / 2, rel=1e-11)
    assert (r["chi_vbs"][:, diag] > 0).all()
    assert abs(r["wf_sgn"]) <= 1.0 + 1e-12
