"""The Fortran driver (rbm_pp_amd/fortran/rbmpp_vmc, main.f90's optimisation
loop over the iso_c_binding module) against the same calls made through the
Python mirror: identical energies and optimised parameters on a 4x4 run
directory in the reference's file formats."""
import os
import subprocess

import numpy as np
import pytest

from rbm_pp_amd import lattice as lt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "rbm_pp_amd", "fortran", "rbmpp_vmc")


def test_fortran_driver_matches_python_mirror(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang at build time)")
    from rbm_pp_amd import api
    W, nspw, K = 64, 20, 16
    spec = lt.j1j2_model(4, alphar=1, J2=0.5)
    spec.nsample, spec.nstep, spec.nav = W * nspw, 3, 2
    spec.delta_tau, spec.s_factor0, spec.s_factor_scale = 0.05, 1e-3, 1.0
    lt.write_run_dir(str(tmp_path), spec, lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.05))
    print(f"fortran driver: 4x4 alphar=1 walkers={W} samples/walker={nspw}", flush=True)
    out = subprocess.run([DRIVER, str(W), "0", str(K)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout[-2000:]
    e_f = np.loadtxt(tmp_path / "Energy_vs_Iteration.txt")

    # the same sequence through the Python mirror
    spec2, p = lt.read_run_dir(str(tmp_path))
    eng = api.Engine(spec2, n_walkers=W, n_samples=nspw, seed=1, recompute_k=K,
                     s_precision=api.S_FP64, params=p)
    shift, wf_list, s0 = 0.0, [0.0] * 10, spec.s_factor0
    gam_av = np.zeros(spec.nv)
    for it in range(1, spec.nstep + 1):
        eng.make_sample(it == 1, shift=shift)
        r = eng.calc_phys_quantity_opt(shift=shift)
        gam, _ = eng.sr_step(eng.gamma(), s_factor0=s0, delta_tau=spec.delta_tau, refine_iters=1)
        var = (r["Etot2"] - r["Etot"] ** 2) / r["Etot"] ** 2
        np.testing.assert_allclose(e_f[it - 1], [it, r["Etot"], var], rtol=1e-12)
        if it > spec.nstep - spec.nav:
            gam_av += gam
        s0 *= spec.s_factor_scale
        wf_list = wf_list[1:] + [r["wf_av"]]
        shift = sum(wf_list) / 10.0 * 1e-3 / spec.N / spec.N
    gam_av /= spec.nav
    po = lt.Params.from_gamma(spec, gam_av)
    # the orbital-derivative bins are summed with LDS atomics, whose order
    # varies between runs: O, S and the solved step agree to rounding, which
    # the SR solve turns into absolute parameter differences of ~1e-14
    _, _, Wr, br = lt.read_w(str(tmp_path / "optimized_W.txt"), 0, spec.alphar, spec.N)
    np.testing.assert_allclose(Wr, po.Wr, rtol=0, atol=1e-13)
    np.testing.assert_allclose(br, po.br, rtol=0, atol=1e-13)
    f = lt.read_zqp_orbital(str(tmp_path / "zqp_orbital_opt.dat"), spec.n_slater)
    np.testing.assert_allclose(f, po.f, rtol=0, atol=1e-13)  # e25.15 in the file
    avg = open(tmp_path / "Average_Energy.txt").read().split()
    assert len(avg) == 10 and np.isfinite([float(t) for t in avg]).all()


def test_fortran_correlation_run_matches_python_mirror(tmp_path):
    """Mode 1 of the driver (the Calc_Correlation measurement run) against the
    same call sequence through the Python mirror: fort.<60000+it> (Chi_sp)
    and fort.<70000+it> (Chi_vbs).  The two read the parameter files with
    different parsers and sums (the mean |f| normalisation), so parameters
    and results agree to rounding (1e-13), not bit for bit."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang at build time)")
    from rbm_pp_amd import api
    W, nspw, K = 64, 10, 16
    spec = lt.j1j2_model(4, alphar=2, alphac=1, J2=0.5)
    spec.nsample, spec.nstep, spec.nav = W * nspw, 1, 2
    lt.write_run_dir(str(tmp_path), spec, lt.random_params(spec, seed_w=3, seed_f=4, sigma_w=0.05))
    print(f"fortran driver, correlation run: 4x4 walkers={W} samples/walker={nspw}", flush=True)
    out = subprocess.run([DRIVER, str(W), "0", str(K), "0", "1"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout[-2000:]
    spec2, p = lt.read_run_dir(str(tmp_path))
    eng = api.Engine(spec2, n_walkers=W, n_samples=nspw, seed=1, recompute_k=K, s_precision=api.S_FP64,
                     params=p)
    eng.make_sample(True, shift=0.0)
    shift = eng.calc_phys_quantity_after_opt(shift=0.0)["wf_av"] * 1e-5
    for _ in range(7):
        eng.make_sample(False, shift=shift)
    nchi = spec.N * (spec.N + 1) // 2
    for it in (1, 2):
        eng.make_sample(False, shift=shift)
        r = eng.measure_correlation(shift=shift)
        sp = np.loadtxt(tmp_path / f"fort.{60000 + it}", comments="#")
        vb = np.loadtxt(tmp_path / f"fort.{70000 + it}", comments="#")
        assert sp.shape == (nchi, 3) and vb.shape == (nchi, 5)
        np.testing.assert_array_equal(sp[:, 0], np.arange(1, nchi + 1))
        np.testing.assert_allclose(sp[:, 1:], r["chi_sp"], rtol=0, atol=1e-13)
        np.testing.assert_allclose(vb[:, 1:], r["chi_vbs"].T, rtol=0, atol=1e-13 * max(1.0, np.abs(vb).max()))
        e = np.loadtxt(tmp_path / "energy_after_opt.txt")
        assert e[it - 1, 1] == pytest.approx(r["Etot"], rel=1e-12)


def test_fortran_driver_against_oracle_loop(tmp_path, oracle_lib):
    """Three SR iterations of the Fortran driver (4x4, 32 walkers) against the
    same iterations restated on the CPU: the oracle's make_sample on each
    walker's Philox stream (the draws the GPU sampler makes), its
    calc_phys_quantity_opt over all samples and its SR step, with the shift
    update of main.f90:441-451.  The energies of every iteration agree to
    1e-10 (the first) and 1e-8 (later ones, whose parameters come from two
    solves of the same system), and the final parameters to 1e-8."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang at build time)")
    W, nspw, K = 32, 12, 16
    spec = lt.j1j2_model(4, alphar=1, J2=0.5)
    spec.nsample, spec.nstep, spec.nav = W * nspw, 3, 1
    spec.delta_tau, spec.s_factor0, spec.s_factor_scale = 0.05, 1e-3, 1.0
    lt.write_run_dir(str(tmp_path), spec, lt.random_params(spec, seed_w=4, seed_f=5, sigma_w=0.05))
    print(f"fortran driver vs oracle: 4x4 walkers={W} samples/walker={nspw}", flush=True)
    out = subprocess.run([DRIVER, str(W), "0", str(K)], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout[-2000:]
    e_f = np.loadtxt(tmp_path / "Energy_vs_Iteration.txt")
    spec2, p = lt.read_run_dir(str(tmp_path))
    orc = oracle_lib.Oracle(spec2, p)
    chains = [(oracle_lib.Rng("philox", 1, stream=w), orc.new_state()) for w in range(W)]
    nfast = [None] * W
    nwarm = max(30, nspw // 100)  # make_sample.f90:66
    shift, wf_list, gam = 0.0, [0.0] * 10, orc.gamma()
    for it in range(1, spec.nstep + 1):
        xs = []
        for w, (rng, st) in enumerate(chains):
            o = orc.make_sample(rng, it == 1, nsample=nspw, nwarmup=nwarm, fixed_k=K, ref_draws=False,
                                shift=shift, state=st, cond_guard=1e4, nfastup=nfast[w])
            nfast[w] = o["nfastup"]
            xs.append(o["x"])
        m = orc.measure(np.concatenate(xs), shift=shift, want_s=True)
        etot = m["Etot"] / m["norm"]
        assert e_f[it - 1, 1] == pytest.approx(etot, rel=1e-10 if it == 1 else 1e-8), it
        rc, gam, _, _ = orc.sr_step(etot, m["norm"], m["gvec"], m["ovec"], m["smat"], gam, s_factor0=1e-3)
        assert rc == 0
        orc.set_gamma(gam)
        wf_list = wf_list[1:] + [m["wf_av"] / (W * nspw)]
        shift = sum(wf_list) / 10.0 * 1e-3 / spec.N / spec.N
    Wc, bc, Wr, br = lt.read_w(str(tmp_path / "optimized_W.txt"), spec.alphac, spec.alphar, spec.N)
    f = lt.read_zqp_orbital(str(tmp_path / "zqp_orbital_opt.dat"), spec.n_slater)
    po = lt.Params.from_gamma(spec, gam)
    np.testing.assert_allclose(Wr, po.Wr, rtol=0, atol=1e-8 * np.abs(po.Wr).max())
    np.testing.assert_allclose(f, po.f, rtol=0, atol=1e-8 * np.abs(po.f).max())
