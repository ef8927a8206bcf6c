"""The sampler's layouts for block-form models on the same walkers and seeds.
The slim layout (the default at N <= 104: Slater_Elm entries looked up from
f and the orbital map in LDS) against the full layout (the global Slater_Elm
table): bit-identical samples, amplitudes and F^-1.  The resident form (rows
[0, 16) of every copy's F^-1 kept in LDS for the whole launch, the rest
streamed) against the slim one: the same configurations and acceptance
counts, amplitudes to rounding, and F^-1 as close to a recompute of the
final configuration (the two group the wave sums of the pending update
differently).  The cases
cover the adaptive and a fixed recompute period (the flush before every
checked recompute and the reload after it), a tight conditioning guard
(recomputes after single accepts) and global spin flips."""
import os

import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu


def _relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _run(form, spec, p, W, ns, K, guard, p_global):
    """Two make_sample calls; the samples, the walkers and the drift of each
    walker's F^-1 from a recompute of its final configuration."""
    from rbm_pp_amd import api
    os.environ["RBMPP_SWEEP_RESIDENT"] = "1" if form == "resident" else "0"
    os.environ["RBMPP_SWEEP_SLIM"] = "0" if form == "full" else "1"
    try:
        eng = api.Engine(spec, params=p, n_walkers=W, n_samples=ns, seed=77, n_warmup=20, recompute_k=K,
                         cond_guard=guard, p_global=p_global)
        eng.make_sample(True, shift=0.0)
        eng.make_sample(False, shift=1e-3)
        x, psi = eng.get_samples()
        wk = eng.get_walkers(with_invm=True)
    finally:
        os.environ.pop("RBMPP_SWEEP_RESIDENT", None)
        os.environ.pop("RBMPP_SWEEP_SLIM", None)
    # InvM_walker(i, j) = InvM_eval(rank[i], rank[j]): the walker keeps its
    # own electron labels, rbmpp_eval numbers them in site order per spin
    ele = eng.walker_electrons()
    rec = eng.eval(wk["x"], with_invm=True)["invm"]
    H = spec.N // 2
    drift = np.zeros(W)
    for w in range(W):
        e = ele[w]
        rank = np.empty(spec.N, np.int64)
        rank[:H] = np.argsort(np.argsort(e[:H]))
        rank[H:] = H + np.argsort(np.argsort(e[H:]))
        drift[w] = max(_relerr(wk["invm"][w, q], rec[w, q][np.ix_(rank, rank)]) for q in range(rec.shape[1]))
    return x, psi, wk, drift


@pytest.mark.parametrize("L,ar,K,guard,p_global", [(10, 4, 0, 1e4, 3e-4), (10, 4, 12, 1e4, 0.02),
                                                   (10, 4, 0, 3.0, 3e-4), (8, 2, 0, 1e4, 3e-4)])
def test_sampler_layouts_agree(L, ar, K, guard, p_global):
    spec = lt.j1j2_model(L, alphar=ar, J2=0.5)
    p = lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.1)
    W, ns = 192, 6
    print(f"resident vs streaming: L={L} ar={ar} K={K} guard={guard} p_global={p_global} W={W}", flush=True)
    xs, ps, ws, ds = _run("slim", spec, p, W, ns, K, guard, p_global)
    xf, pf, wf, _ = _run("full", spec, p, W, ns, K, guard, p_global)
    assert np.array_equal(xf, xs)
    assert np.array_equal(pf, ps)
    assert np.array_equal(wf["invm"], ws["invm"])
    assert np.array_equal(wf["naccept"], ws["naccept"])
    xr, pr, wr, dr = _run("resident", spec, p, W, ns, K, guard, p_global)
    assert np.array_equal(xr, xs)
    assert np.array_equal(wr["naccept"], ws["naccept"])
    assert np.array_equal(wr["x"], ws["x"])
    # the two fast-update chains round differently (up to 100 accepts
    # between recomputes): amplitudes agree to 1e-7, and each form's F^-1
    # stays as close to a recompute as the other's
    assert _relerr(pr, ps) < 1e-7
    assert _relerr(wr["pf"], ws["pf"]) < 1e-7
    print(f"F^-1 drift from a recompute: resident {dr.max():.2e}, slim {ds.max():.2e}", flush=True)
    assert dr.max() <= max(1e-7, 4 * ds.max())
