"""The uneven multi-lattice batch: lattices 6x6 ... 14x14 at J2/J1 in
{0.4, 0.5, 0.55}, RBM alpha = 2, 2x2-sublattice pair product, C4v + parity,
walkers apportioned by N^2, sampled by one persistent launch
(rbmpp_batch_make_sample) and measured / optimised per lattice.  The batch
reproduces each lattice's own make_sample bit for bit (the same Philox
streams), and the O_k of every lattice size match central finite
differences of the GPU amplitude."""
import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

SIZES = [6, 8, 10, 12, 14]


def _configs(N, n, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((n, N), np.int32)
    for k in range(n):
        x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
        rng.shuffle(x)
        out[k] = x
    return out


def _engines(api, total, nspw, J2s=(0.5,), seed=7):
    specs = [(L, J2) for L in SIZES for J2 in J2s]
    n2 = np.array([L ** 4 for L, _ in specs], float)
    walkers = np.maximum(1, np.round(total * n2 / n2.sum())).astype(int)
    out = []
    for (L, J2), W in zip(specs, walkers):
        spec = lt.j1j2_model(L, alphar=2, J2=J2)
        p = lt.random_params(spec, seed_w=L, seed_f=L + 50, sigma_w=0.05)
        out.append(api.Engine(spec, params=p, n_walkers=int(W), n_samples=nspw, seed=seed + L, n_warmup=2,
                              recompute_k=0))
    return out


def test_batch_reproduces_separate_launches():
    from rbm_pp_amd import api
    A = _engines(api, 96, 3)
    B = _engines(api, 96, 3)
    print("batch walkers per lattice:", [e.W for e in A], flush=True)
    batch = api.Batch(A)
    for first, shift in ((True, 0.0), (False, 1e-3)):
        nb = batch.make_sample([first] * len(A), [shift] * len(A))
        for k, e in enumerate(B):
            ns = e.make_sample(first, shift=shift)
            assert nb[k] == ns, (k, nb[k], ns)
            xa, pa = A[k].get_samples()
            xb, pb = e.get_samples()
            assert np.array_equal(xa, xb), k
            np.testing.assert_array_equal(pa, pb)
    # each lattice measures and steps on its own after the batch
    for e in A:
        e.calc_phys_quantity_opt(shift=1e-3)
        e.sr_step(e.gamma(), s_factor0=1e-3)


def test_batch_rejects_unsupported_projection():
    from rbm_pp_amd import api
    spec = lt.j1j2_model(6, alphar=2, J2=0.5)
    e = api.Engine(spec, params=lt.random_params(spec), n_walkers=2, n_samples=1)
    e.set_projection(api.PROJ_TRANSVERSE)
    with pytest.raises(api.RbmppError):
        api.Batch([e])


@pytest.mark.parametrize("L", SIZES)
def test_adjoint_ok_against_finite_differences_every_size(L):
    """O_k = d psi / d gamma_k from the adjoint kernels against central
    differences (h = 1e-6) of the GPU amplitude on two configurations;
    above N = 104 on 160 parameters drawn at random (the others are the same
    code).  Relative 2e-8 of each configuration's largest component."""
    from rbm_pp_amd import api
    spec = lt.j1j2_model(L, alphar=2, J2=0.5)
    p = lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.1)
    eng = api.Engine(spec, params=p, n_walkers=1, n_samples=1)
    xs = _configs(spec.N, 2, seed=L)
    O = eng.local(xs, want_o=True)["oloc"]
    g0 = p.gamma()
    ks = np.arange(spec.nv) if spec.N <= 100 else np.sort(np.random.default_rng(L).choice(spec.nv, 160, False))
    h = 1e-6
    fd = np.zeros((len(xs), len(ks)), np.complex128)
    for c, k in enumerate(ks):
        gp = g0.copy(); gp[k] += h
        gm = g0.copy(); gm[k] -= h
        eng.set_gamma(gp)
        a = eng.eval(xs)["psi"][:, 0]
        eng.set_gamma(gm)
        b = eng.eval(xs)["psi"][:, 0]
        fd[:, c] = (a - b) / (2 * h)
    scale = np.abs(O).max(axis=1, keepdims=True)
    err = (np.abs(O[:, ks] - fd) / scale).max()
    assert err < 2e-8, err
