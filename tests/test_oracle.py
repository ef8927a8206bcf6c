"""Pins the CPU oracle (oracle/rbmpp_oracle.c) before it is trusted as the
checker for the HIP path: reference golden vectors, the compiled reference
routines when present, closed forms, brute force and exact enumeration."""
import itertools
import os

import numpy as np
import pytest

import ed
from oracle import refharness as RH
from rbm_pp_amd import lattice as lt

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _model(L=4, alphar=1, alphac=0, J2=0.5, seed=1, sigma_w=0.05, **kw):
    spec = lt.j1j2_model(L, alphar=alphar, alphac=alphac, J2=J2, **kw)
    return spec, lt.random_params(spec, seed_w=seed, seed_f=seed + 100, sigma_w=sigma_w)


def _rand_x(rng, N):
    x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
    rng.shuffle(x)
    return x


# --------------------------------------------------------------------- RNGs
def test_mt19937_matches_reference_stream(oracle_lib):
    g = np.load(os.path.join(GOLD, "ref_mt19937.npz"))
    for key, seed in (("seed1", 1), ("seed4357", 4357), ("seed123456789", 123456789)):
        r = oracle_lib.Rng("mt", seed)
        assert np.array_equal(r.u01(2000), g[key])


def test_philox_known_answers(oracle_lib):
    # Random123 known-answer vectors for philox4x32-10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert tuple(int(v) for v in oracle_lib.philox4x32_10(ctr, key)) == want


# ---------------------------------------------------------------- Pfaffian
def _pf_brute(A):
    n = A.shape[0]
    if n == 0:
        return 1.0
    tot = 0.0
    for j in range(1, n):
        rest = [k for k in range(n) if k not in (0, j)]
        tot += (-1) ** (j - 1) * A[0, j] * _pf_brute(A[np.ix_(rest, rest)])
    return tot


def test_pfaffian_brute_force_and_det(oracle_lib):
    rng = np.random.default_rng(3)
    for n in (2, 4, 6, 8):
        for _ in range(5):
            B = rng.normal(size=(n, n)); A = B - B.T
            pf = oracle_lib.pfaffian(A)
            assert pf == pytest.approx(_pf_brute(A), rel=1e-12)
            assert pf * pf == pytest.approx(np.linalg.det(A), rel=1e-10)
    B = rng.normal(size=(4, 4)); A = B - B.T
    closed = A[0, 1] * A[2, 3] - A[0, 2] * A[1, 3] + A[0, 3] * A[1, 2]
    assert oracle_lib.pfaffian(A) == pytest.approx(closed, rel=1e-13)
    # Pf [[0, M], [-M^T, 0]] = (-1)^{n(n-1)/2} det M
    for n in (3, 5):
        M = rng.normal(size=(n, n))
        A = np.block([[np.zeros((n, n)), M], [-M.T, np.zeros((n, n))]])
        assert oracle_lib.pfaffian(A) == pytest.approx((-1) ** (n * (n - 1) // 2) * np.linalg.det(M), rel=1e-11)
    for n in (24, 40):
        B = rng.normal(size=(n, n)); A = B - B.T
        pf = oracle_lib.pfaffian(A)
        sign, logdet = np.linalg.slogdet(A)
        assert 2 * np.log(abs(pf)) == pytest.approx(logdet, rel=1e-12)


# ------------------------------------------------- reference golden vectors
@pytest.mark.parametrize("name", ["ref_4x4_a1c0", "ref_4x4_a2c1", "ref_6x6_a2c0", "ref_4x4_a2c1_tc"])
def test_oracle_matches_reference_goldens(oracle_lib, name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    spec = lt.j1j2_model(int(g["L"]), alphar=int(g["alphar"]), alphac=int(g["alphac"]), J2=float(g["J2"]))
    p = lt.Params(Wc=g["Wc"], bc=g["bc"], Wr=g["Wr"], br=g["br"], f=g["f"])
    orc = oracle_lib.Oracle(spec, p)
    if "projection" in g.files:
        orc.set_projection(int(g["projection"]))
    assert np.array_equal(orc.slater_elm()[0], g["slater_elm_q0"])
    assert np.array_equal(orc.sp_weights(), g["sp"])
    for k in range(g["x"].shape[0]):
        st = orc.eval(g["x"][k])
        assert np.array_equal(st["thetar"], g["thetar"][k])
        assert np.array_equal(st["thetac"], g["thetac"][k])
        assert np.array_equal(st["ele_idx"], g["ele_idx"][k])
        assert np.array_equal(st["ele_cfg"], g["ele_cfg"][k])
        assert np.array_equal(st["sgn"], g["sgn"][k])
        assert np.array_equal(st["pf"], g["pf_in"][k])
        np.testing.assert_allclose(st["psi"], g["psi"][k], rtol=1e-13, atol=0)
        assert abs(orc.eloc() - g["eloc"][k]) <= 1e-12 * abs(g["eloc"][k])
        opp = orc.oloc()[spec.nv_rbm:]
        np.testing.assert_allclose(opp, g["opp"][k], rtol=0, atol=1e-13 * np.abs(g["opp"][k]).max())
        ratio, pfn, psin = orc.ratio_exchange(int(g["ex_i1"][k]), int(g["ex_i2"][k]))
        np.testing.assert_allclose(pfn, g["pf_two"][k], rtol=1e-12)
        np.testing.assert_allclose(psin, g["psi_two"][k], rtol=1e-12)
        if "invm_two" in g.files and k < g["invm_two"].shape[0]:
            orc.accept_exchange(int(g["ex_i1"][k]), int(g["ex_i2"][k]))
            inv = orc.state()["invm"]
            np.testing.assert_allclose(inv, g["invm_two"][k], rtol=0, atol=1e-13 * np.abs(inv).max())


@pytest.mark.parametrize("projection", [0, 1])
def test_oracle_matches_live_reference(oracle_lib, projection):
    if not RH.available(projection):
        pytest.skip("reference routines not built here")
    spec, p = _model(L=6, alphar=2, alphac=1, seed=7)
    orc = oracle_lib.Oracle(spec, p)
    orc.set_projection(projection)
    ref = RH.Ref(spec, p, projection=projection)
    assert np.array_equal(orc.sp_weights(), ref.sp())
    assert np.array_equal(orc.slater_elm(), ref.slater_elm())
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = _rand_x(rng, spec.N)
        st = orc.eval(x)
        assert orc.eloc() == ref.eloc(st)
        i1 = 0
        i2 = int(np.nonzero(x != x[0])[0][0])
        _, pfn, _ = orc.ratio_exchange(i1, i2)
        tmp = st["ele_idx"].copy(); msa, msb = st["ele_cfg"][i1], st["ele_cfg"][i2]
        tmp[msa] = i2; tmp[msb] = i1
        assert np.array_equal(pfn, ref.pfm_two(msa, msb, tmp, st["invm"], st["pf"]))
    for t in range(4):
        for i in range(spec.N):
            from oracle import lib
            assert lib().om_bond(spec.L, t, i) == ref.bond(t, i)


# ------------------------------------------------ algebraic / physical pins
def test_inverse_and_pfaffian_of_each_copy(oracle_lib):
    spec, p = _model(L=4, alphar=2, alphac=1)
    orc = oracle_lib.Oracle(spec, p)
    se = orc.slater_elm()
    rng = np.random.default_rng(1)
    N, np_ = spec.N, spec.N // 2
    for _ in range(4):
        st = orc.eval(_rand_x(rng, N))
        r = st["ele_idx"] + N * (np.arange(N) >= np_)
        for q in range(spec.C):
            X = -se[q][np.ix_(r, r)]
            assert np.abs(X + X.T).max() == 0.0
            inv = st["invm"][q]          # InvM(i,j) = Xinv(j,i)
            np.testing.assert_allclose(X @ inv.T, np.eye(N), atol=1e-11)
            assert st["pf"][q] ** 2 == pytest.approx(np.linalg.det(X), rel=1e-9)


def test_fast_update_matches_recompute(oracle_lib):
    spec, p = _model(L=6, alphar=2, alphac=1)
    orc = oracle_lib.Oracle(spec, p)
    rng = np.random.default_rng(2)
    x = _rand_x(rng, spec.N)
    orc.eval(x)
    for _ in range(25):
        st = orc.state()
        x = st["x"]
        i1 = int(rng.integers(spec.N))
        i2 = int(rng.choice(np.nonzero(x != x[i1])[0]))
        _, _, psin = orc.ratio_exchange(i1, i2)
        x2 = x.copy(); x2[i1] *= -1; x2[i2] *= -1
        direct = orc.__class__(spec, p).eval(x2)["psi"][0]
        assert abs(psin[0] - direct) <= 1e-9 * abs(direct)
        orc.accept_exchange(i1, i2)
    drift = orc.recompute()
    assert drift.max() < 1e-7


def test_local_energy_exact_enumeration_4x4(oracle_lib):
    """sum_x |psi|^2 E_loc / sum |psi|^2 == <psi|H|psi>/<psi|psi> with H built
    independently from the bond list; the gauged H has the exact 4x4 J1-J2
    ground energy (known value -0.52862/site at J2/J1 = 0.5 in S.S units)."""
    spec, p = _model(L=4, alphar=1, alphac=1)
    orc = oracle_lib.Oracle(spec, p)
    states, Hg = ed.hamiltonian(4, 1.0, 0.5, gauge=True)
    m = orc.measure(states.astype(np.int32), shift=0.0, want_s=False)
    psi = m["psi"]
    eloc = m["eloc"] * np.abs(psi) / psi
    e_mc = (np.abs(psi) ** 2 * eloc).sum() / (np.abs(psi) ** 2).sum()
    e_h = (np.conj(psi) @ (Hg @ psi)).real / (np.abs(psi) ** 2).sum()
    assert abs(e_mc.imag) < 1e-10
    assert e_mc.real == pytest.approx(e_h, abs=1e-12)
    import scipy.sparse.linalg as sla
    e0g = sla.eigsh(Hg, k=1, which="SA")[0][0]
    assert e0g == pytest.approx(ed.ground_energy(4, 1.0, 0.5), abs=1e-9)
    assert e0g / 4 / 16 == pytest.approx(-0.528620, abs=2e-6)
    assert e_h >= e0g


# ------------------------------------------------ correlation run (Calc_Correlation)
def _cc_model(g):
    spec = lt.j1j2_model(int(g["L"]), alphar=int(g["alphar"]), alphac=int(g["alphac"]), J2=float(g["J2"]))
    return spec, lt.Params(Wc=g["Wc"], bc=g["bc"], Wr=g["Wr"], br=g["br"], f=g["f"])


def test_correlation_matches_reference_goldens(oracle_lib):
    """SdotS per bond and Chi_loc_x per pair, bit-exact against the
    Calc_Correlation tree's own routines."""
    g = np.load(os.path.join(GOLD, "ref_4x4_a2c1_cc.npz"))
    spec, p = _cc_model(g)
    orc = oracle_lib.Oracle(spec, p)
    for k in range(g["x"].shape[0]):
        orc.eval(g["x"][k])
        e, sd = orc.bond_sdots()
        assert e == g["eloc"][k]
        assert np.array_equal(sd, g["sdots"][k])
        assert np.array_equal(orc.local_correlation(), g["chi_loc"][k])


def test_correlation_matches_live_reference(oracle_lib):
    if not RH.available(tree="cc"):
        pytest.skip("reference routines not built here")
    spec, p = _model(L=6, alphar=2, alphac=1, seed=8)
    orc = oracle_lib.Oracle(spec, p)
    ref = RH.Ref(spec, p, tree="cc")
    rng = np.random.default_rng(3)
    for _ in range(2):
        st = orc.eval(_rand_x(rng, spec.N))
        e, sd = orc.bond_sdots()
        e_r, sd_r = ref.eloc_sdots(st)
        assert e == e_r and np.array_equal(sd, sd_r)
        assert np.array_equal(orc.local_correlation(), ref.local_correlation(st))


def test_correlation_exact_enumeration_4x4(oracle_lib):
    """sum_x conj(psi) Chi_loc_x / sum |psi|^2 equals <psi|O|psi>/<psi|psi>
    for sigma^z sigma^z and the gauged exchange of every pair, with O built
    from the state list alone; the bond SdotS sum to E_loc."""
    spec, p = _model(L=4, alphar=1, alphac=1)
    orc = oracle_lib.Oracle(spec, p)
    states, ops = ed.pair_operators(4, gauge=True)
    nchi = len(ops)
    psi = np.zeros(len(states), complex)
    acc = np.zeros((nchi, 2), complex)
    for k, x in enumerate(states.astype(np.int32)):
        st = orc.eval(x)
        psi[k] = st["psi"][0]
        acc += np.conj(psi[k]) * orc.local_correlation()
        if k % 997 == 0:
            e, sd = orc.bond_sdots()
            J = np.array([1.0, 1.0, 0.5, 0.5])
            assert (J[:, None] * sd).sum() == pytest.approx(e, rel=1e-13, abs=1e-13 * abs(psi[k]))
    nrm = (np.abs(psi) ** 2).sum()
    for i0, (i1, i2, zz, ex) in enumerate(ops):
        assert acc[i0, 0].real / nrm == pytest.approx((np.abs(psi) ** 2 * zz).sum() / nrm, abs=1e-12)
        assert acc[i0, 1].real / nrm == pytest.approx((np.conj(psi) @ (ex @ psi)).real / nrm, abs=1e-12)
        assert abs(acc[i0, 1].imag) / nrm < 1e-12


def test_correlation_sums(oracle_lib):
    """om_measure_correlation against the sums of the reference's own
    per-sample values (goldens) in the reference's accumulation
    (Calc_Correlation/calc_phys_quantity.f90:318-338), and the closed forms:
    Chi_sp(.,ii) = norm, sum over i<j of Chi_sp(1,.) = -N norm / 2 (Sz = 0)."""
    g = np.load(os.path.join(GOLD, "ref_4x4_a2c1_cc.npz"))
    spec, p = _cc_model(g)
    orc = oracle_lib.Oracle(spec, p)
    N, shift = spec.N, 0.3
    xs = g["x"]
    got = orc.measure_correlation(xs, shift=shift)
    nchi = N * (N + 1) // 2
    sp_ref = np.zeros((nchi, 2)); vb_ref = np.zeros((4, nchi)); norm = 0.0; wsg = 0.0
    iu = [(i1, i2) for i2 in range(N) for i1 in range(i2 + 1)]
    for k, x in enumerate(xs):
        psi = orc.eval(x)["psi"][0]
        rw = 1.0 / (abs(psi) + shift)
        norm += (abs(psi) * rw) ** 2
        wsg += psi / abs(psi)
        sd = g["sdots"][k] * rw
        for t in range(4):
            vb_ref[t] += np.array([(np.conj(sd[t, a]) * sd[t, b]).real for a, b in iu])
        sp_ref += rw * rw * (np.conj(psi) * g["chi_loc"][k]).real
    assert got["norm"] == pytest.approx(norm, rel=1e-14)
    assert got["wf_sgn"] == pytest.approx(wsg, rel=1e-14)
    np.testing.assert_allclose(got["chi_sp"], sp_ref, rtol=1e-13, atol=1e-15 * norm)
    np.testing.assert_allclose(got["chi_vbs"], vb_ref, rtol=1e-13, atol=1e-15 * np.abs(vb_ref).max())
    diag = [i2 * (i2 + 1) // 2 + i2 for i2 in range(N)]
    np.testing.assert_allclose(got["chi_sp"][diag], norm, rtol=1e-13)
    off = np.ones(nchi, bool); off[diag] = False
    assert got["chi_sp"][off, 0].sum() == pytest.approx(-N * norm / 2, rel=1e-12)
    e = orc.measure_energy(xs, shift=shift)
    assert (got["Etot"], got["Etot2"], got["wf_av"]) == (e["Etot"], e["Etot2"], e["wf_av"])


def test_oloc_central_finite_differences(oracle_lib):
    """O_loc(k) = d psi / d gamma_k against central differences, h=1e-6."""
    spec, p = _model(L=4, alphar=1, alphac=1, seed=4)
    orc = oracle_lib.Oracle(spec, p)
    gam0 = orc.gamma()
    rng = np.random.default_rng(9)
    x = _rand_x(rng, spec.N)
    orc.eval(x)
    ol = orc.oloc()
    h = 1e-6
    fd = np.zeros(spec.nv, np.complex128)
    probe = oracle_lib.Oracle(spec)
    for k in range(spec.nv):
        gp = gam0.copy(); gp[k] += h
        gm = gam0.copy(); gm[k] -= h
        probe.set_gamma(gp); a = probe.eval(x)["psi"][0]
        probe.set_gamma(gm); b = probe.eval(x)["psi"][0]
        fd[k] = (a - b) / (2 * h)
    scale = np.abs(ol).max()
    np.testing.assert_allclose(ol, fd, rtol=0, atol=2e-8 * scale)
    # f_ii (displacement-zero orbitals) never enter X: exactly zero derivative
    fii = {int(spec.orb[i, i]) for i in range(spec.N)}
    assert all(ol[spec.nv_rbm + o] == 0 for o in fii)


def test_measure_and_sr_invariants(oracle_lib):
    spec, p = _model(L=4, alphar=2, alphac=0)
    orc = oracle_lib.Oracle(spec, p)
    rng = np.random.default_rng(5)
    xs = np.array([_rand_x(rng, spec.N) for _ in range(200)])
    m = orc.measure(xs, shift=1e-3, want_s=True, want_ostore=True)
    O = m["ostore"]
    # entrywise 1e-12, and for entries that cancel, 1e-14 of the largest
    ref = O.T @ O
    np.testing.assert_allclose(m["smat"], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    nv, norm = spec.nv, m["norm"]
    ovec = m["ovec"] / norm
    S = m["smat"] / norm - np.outer(ovec, ovec)
    w = np.linalg.eigvalsh(S)
    assert w.min() > -1e-12 * w.max()
    etot = m["Etot"] / norm
    gam = orc.gamma()
    rc, gam_new, rep, x = orc.sr_step(etot, norm, m["gvec"], m["ovec"], m["smat"], gam,
                                      s_factor0=1e-3)
    assert rc == 0
    # mean |f| renormalised to 0.25 (SR_pdposv.f90:414)
    assert np.mean(np.abs(gam_new[spec.nv_rbm:])) == pytest.approx(0.25, abs=1e-12)
    # the solution solves the shifted, rescaled system: rebuild it independently
    x_np, lam = _sr_solution_numpy(spec, m, 1e-3)
    assert rep[0] == pytest.approx(lam, rel=1e-9)
    np.testing.assert_allclose(x, x_np, rtol=1e-8, atol=1e-12 * np.abs(x_np).max())


def _sr_solution_numpy(spec, m, s_factor0):
    """SR_pdposv.f90:4-323 rebuilt with numpy: centred S and g, PP block
    scaling by sqrt(mean positive RBM diagonal / mean positive PP diagonal),
    relative diagonal shift max(1e-4, s_factor0) of the mean positive
    diagonal, dense solve, trust region lambda = sqrt(x^T A x) 5/L, scaling
    undone.  Returns (x, lambda)."""
    nv, nvr, norm = spec.nv, spec.nv_rbm, m["norm"]
    ovec = m["ovec"] / norm
    etot = m["Etot"] / norm
    S = m["smat"] / norm - np.outer(ovec, ovec)
    g = m["gvec"] / norm - 2 * etot * ovec
    d = np.diag(S)
    scal = np.sqrt(d[:nvr][d[:nvr] > 0].mean() / d[nvr:][d[nvr:] > 0].mean())
    D = np.ones(nv); D[nvr:] = scal
    Ss = S * np.outer(D, D); gs = g * D
    Sav = np.diag(Ss)[np.diag(Ss) > 0].mean()
    Ss[np.diag_indices(nv)] += Sav * max(1e-4, s_factor0)
    xs_ = np.linalg.solve(Ss, gs)
    lam = np.sqrt(xs_ @ Ss @ xs_) * 5 / spec.L
    if lam > 1:
        xs_ = xs_ / lam
    return xs_ * D, lam


def _sr_update_numpy(gam0, x, delta_tau, max_change, nsl, slater_abs_av=0.25):
    """SR_pdposv.f90:329-366 rebuilt from the solution x: the max-change clip
    scales the whole step so that its largest component is max_change, then
    the pair amplitudes are renormalised to mean |f| = slater_abs_av."""
    step = delta_tau * np.asarray(x)
    dmax = np.abs(step).max()
    if dmax > max_change:
        step = step * (max_change / dmax)
    gam = np.asarray(gam0) - step
    f = gam[len(gam) - nsl:]
    gam[len(gam) - nsl:] = f * (slater_abs_av / np.abs(f).mean())
    return gam, dmax, step


@pytest.mark.parametrize("delta_tau,clipped", [(1.0, True), (0.5, True), (1e-4, False)])
def test_sr_step_clip_and_gamma_update(oracle_lib, delta_tau, clipped):
    """Both branches of the max-change clip of SR_pdposv.f90:329-340 and the
    f renormalisation of :342-354, against an independent numpy rebuild from
    the solution x (which test_measure_and_sr_invariants pins by a numpy
    solve of the scaled, shifted system)."""
    spec, p = _model(L=4, alphar=2, alphac=0)
    orc = oracle_lib.Oracle(spec, p)
    rng = np.random.default_rng(15)
    xs = np.array([_rand_x(rng, spec.N) for _ in range(200)])
    m = orc.measure(xs, shift=1e-3, want_s=True)
    etot = m["Etot"] / m["norm"]
    gam0 = orc.gamma()
    max_change = 0.04
    rc, gam, rep, x = orc.sr_step(etot, m["norm"], m["gvec"], m["ovec"], m["smat"], gam0,
                                  s_factor0=1e-3, delta_tau=delta_tau, max_change=max_change)
    assert rc == 0
    x_np, lam = _sr_solution_numpy(spec, m, 1e-3)
    np.testing.assert_allclose(x, x_np, rtol=1e-8, atol=1e-12 * np.abs(x_np).max())
    want, dmax, step = _sr_update_numpy(gam0, x, delta_tau, max_change, spec.n_slater)
    assert (dmax > max_change) == clipped
    assert rep[1] == pytest.approx(dmax, rel=1e-15)  # delW_max
    np.testing.assert_allclose(gam, want, rtol=1e-14, atol=1e-15)
    if clipped:
        # the clipped step's largest component is max_change exactly and the
        # step stays parallel to x
        assert np.abs(step).max() == pytest.approx(max_change, rel=1e-14)
        k = int(np.argmax(np.abs(x)))
        np.testing.assert_allclose(step, x * (step[k] / x[k]), rtol=1e-13, atol=1e-18)
    else:
        np.testing.assert_allclose(step, delta_tau * x, rtol=1e-15)
    # the rbm part of gamma moves by the step alone; f is renormalised
    nvr = spec.nv_rbm
    np.testing.assert_allclose(gam[:nvr], gam0[:nvr] - step[:nvr], rtol=1e-14, atol=1e-16)
    assert np.abs(gam[nvr:]).mean() == pytest.approx(0.25, rel=1e-14)
    # a planted mistake fails: clipping only the largest component
    bad = gam0 - delta_tau * x
    kmax = int(np.argmax(np.abs(delta_tau * x)))
    bad[kmax] = gam0[kmax] - np.sign(x[kmax]) * max_change
    if clipped:
        assert np.abs(bad[:nvr] - gam[:nvr]).max() > 1e-6


def test_sampler_energy_4x4_matches_enumeration(oracle_lib):
    """VMC estimate from the reference-ordered sampler (mt.f stream) agrees
    with the exact expectation of the same wave function."""
    spec, p = _model(L=4, alphar=1, alphac=0, sigma_w=0.1)
    orc = oracle_lib.Oracle(spec, p)
    states, Hg = ed.hamiltonian(4, 1.0, 0.5, gauge=True)
    m = orc.measure(states.astype(np.int32), shift=0.0, want_s=False)
    psi = m["psi"]
    e_exact = (np.conj(psi) @ (Hg @ psi)).real / (np.abs(psi) ** 2).sum()
    rng = oracle_lib.Rng("mt", 12345)
    s = orc.make_sample(rng, True, nsample=3000, nwarmup=50)
    mm = orc.measure_energy(s["x"], shift=0.0)
    e = mm["Etot"] / mm["norm"]
    # crude error bar: E_loc spread / sqrt(n/tau), tau ~ 2 sweeps
    m2 = orc.measure(s["x"], shift=0.0, want_s=False)
    el = (m2["eloc"] * np.abs(m2["psi"]) / m2["psi"]).real
    err = el.std() / np.sqrt(len(el) / 4.0)
    assert abs(e - e_exact) < 5 * err
    assert 0.0 < s["naccept"] / (3000 * spec.N) < 1.0


def test_file_formats_round_trip(tmp_path):
    spec, p = _model(L=4, alphar=2, alphac=1)
    spec.nsample, spec.nstep, spec.nav = 1000, 10, 2
    lt.write_run_dir(str(tmp_path), spec, p)
    s2, p2 = lt.read_run_dir(str(tmp_path))
    assert np.array_equal(s2.w_map, spec.w_map) and np.array_equal(s2.orb, spec.orb)
    assert s2.n_slater == spec.n_slater and s2.nv == spec.nv
    np.testing.assert_allclose(p2.f, p.f, rtol=1e-13)
    np.testing.assert_allclose(p2.Wr, p.Wr, atol=1e-15)
    np.testing.assert_allclose(p2.Wc, p.Wc, atol=1e-15)
    g = p.gamma()
    np.testing.assert_array_equal(lt.Params.from_gamma(spec, g).gamma(), g)


def test_fast_update_near_singular_chain_and_guard(oracle_lib):
    """The reference's fast update (pfupdate_two.f90) run through a chain that
    passes close to a singular configuration of one projection copy: 10x10,
    alpha=4, walker 354 of Philox key 1, second make_sample call.  Without a
    guard the chain's Pf of that copy loses all accuracy (relative error of
    the stored amplitude > 1e10 against a fresh evaluation); with a recompute
    whenever a copy's Pf ratio leaves (1e-4, 1e4) every stored amplitude
    stays within 1e-6 of the fresh evaluation."""
    spec = lt.j1j2_model(10, alphar=4, J2=0.5)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    orc = oracle_lib.Oracle(spec, p)
    worst = {}
    for guard in (0.0, 1e4):
        rng = oracle_lib.Rng("philox", 1, stream=354)
        st = orc.new_state()
        orc.make_sample(rng, True, nsample=49, nwarmup=30, fixed_k=32, ref_draws=False, state=st,
                        cond_guard=guard)
        out = orc.make_sample(rng, False, nsample=49, nwarmup=30, fixed_k=32, ref_draws=False,
                              state=st, cond_guard=guard)
        err = 0.0
        for k in range(49):
            ex = orc.eval(out["x"][k])["psi"][0]
            err = max(err, abs(out["psi"][k] - ex) / abs(ex))
        worst[guard] = err
    assert worst[0.0] > 1e10
    assert worst[1e4] < 1e-6


# ------------------------------------------- run-directory files (parameter.f90)
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rbm_pp_amd", "fortran",
                      "rbmpp_vmc")


@pytest.mark.parametrize("L,ar,ac", [(4, 2, 1), (6, 2, 0)])
def test_run_directory_files_against_reference_reader_and_writer(tmp_path, L, ar, ac):
    """A run directory written by rbm_pp_amd.lattice is read by the
    reference's own initialize_parameters (parameter.f90:5-144, compiled from
    the reference sources) to exactly the parameters the Python reader gives;
    the Fortran driver's reader and writer (its mode 2) and the reference's
    initialize_parameters + write_optimized_parameters (parameter.f90:238-276)
    produce byte-identical optimized_W.txt and zqp_orbital_opt.dat from the
    same files."""
    import shutil
    import subprocess
    if not RH.available(0):
        pytest.skip("reference routines not built here")
    spec, p = _model(L=L, alphar=ar, alphac=ac, seed=11)
    spec.nsample, spec.nstep, spec.nav = 64, 1, 1
    a, b = tmp_path / "driver", tmp_path / "reference"
    lt.write_run_dir(str(a), spec, p)
    shutil.copytree(a, b)
    spec2, p_py = lt.read_run_dir(str(a))
    ref = RH.Ref(spec, p)
    N = spec.N
    wc = np.zeros(2 * max(ac, 1) * N); bc = np.zeros(2 * max(ac, 1)); wr = np.zeros(max(ar, 1) * N)
    br = np.zeros(max(ar, 1)); f = np.zeros(spec.n_slater); orb = np.zeros(N * N, np.int32)
    import ctypes as C
    ref.L.ref_get_params.argtypes = [RH.dp] * 5 + [RH.ip]
    cwd = os.getcwd()
    try:
        os.chdir(b)
        ref.L.ref_sgrnd(1)
        ref.L.ref_initialize_parameters()
        ref.L.ref_get_params(wc, bc, wr, br, f, orb)
        ref.L.ref_write_optimized_parameters()
    finally:
        os.chdir(cwd)
    assert np.array_equal(orb.reshape(N, N), spec2.orb)
    np.testing.assert_array_equal(wr.reshape(-1, N)[:ar], p_py.Wr)
    np.testing.assert_array_equal(br[:ar], p_py.br)
    if ac:
        np.testing.assert_array_equal(wc.reshape(ac, N, 2)[..., 0] + 1j * wc.reshape(ac, N, 2)[..., 1], p_py.Wc)
        np.testing.assert_array_equal(bc.reshape(ac, 2)[:, 0] + 1j * bc.reshape(ac, 2)[:, 1], p_py.bc)
    np.testing.assert_array_equal(f, p_py.f)
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built")
    out = subprocess.run([DRIVER, "1", "0", "0", "0", "2"], cwd=a, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr + out.stdout
    for name in ("optimized_W.txt", "zqp_orbital_opt.dat"):
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
