"""rbmpp_local_noalloc (device staging in the library's fixed array, no
allocation) against rbmpp_local (the allocating entry point, same estimator
kernels) through the C ABI: two calls in a row, the second with fewer
configurations, relative 1e-10; and the size limit."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

RTOL = 1e-10


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


# (L, alphar, J2, projection): block-form estimator with per-workgroup F^-1
# scratch; the skew estimator without scratch; N > 104 with the recompute
# matrix in scratch
CASES = [(10, 4, 0.5, "normal"), (6, 2, 0.5, "normal_exact"), (16, 16, 0.55, "normal")]


@pytest.mark.parametrize("L,ar,J2,proj", CASES)
def test_local_noalloc_matches_local(L, ar, J2, proj):
    from rbm_pp_amd import api
    spec = lt.j1j2_model(L, alphar=ar, J2=J2)
    p = lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.05)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    if proj == "normal_exact":
        eng.set_projection(api.PROJ_NORMAL_EXACT)
    n1 = 200 if spec.N <= 100 else 80
    for n, seed in ((n1, 11), (n1 // 3, 12)):
        xs = _configs(spec.N, n, seed)
        got = eng.local_noalloc(xs)
        ref = eng.local(xs, want_o=False)
        assert got["eloc"].shape == (n,)
        assert _relerr(got["eloc"], ref["eloc"]) < RTOL, (n, seed)
        assert _relerr(got["psi"], ref["psi"]) < RTOL, (n, seed)


def test_local_noalloc_rejects_sizes_beyond_the_array():
    from rbm_pp_amd import api
    spec = lt.j1j2_model(10, alphar=4, J2=0.5)
    p = lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.05)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    # more configurations than the whole 64 MiB array could stage
    n = 64 * 2**20 // (5 * spec.N + 44) + 1
    xs = np.tile(_configs(spec.N, 1, 5), (n, 1))
    with pytest.raises(api.RbmppError) as e:
        eng.local_noalloc(xs)
    assert e.value.code == -1 and "exceeds the staging limit" in str(e.value)
    # the handle stays usable
    small = _configs(spec.N, 4, 6)
    assert _relerr(eng.local_noalloc(small)["eloc"], eng.local(small, want_o=False)["eloc"]) < RTOL
