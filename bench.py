"""Headline benchmark: MC samples/s and SR-step seconds of the VMC hot path
on the 10x10 J1-J2 model (J2/J1 = 0.5, RBM alpha=4, C4v x spin-parity
projection: 16 Pfaffian copies of 100x100, 4096 walkers, 2x10^5 samples per
SR step), mixed-precision S (FP32 products, FP64 centring / solve) against
the plain-FP64 S on the same samples.

One "step" = make_sample (30 thermalisation sweeps + 49 stored sweeps per
walker) + measurement (E_loc, O_k) + SR step, i.e. one SR iteration of
main.f90:311-455.  Inputs are synthetic (random W, b ~ N(0, 0.01^2), f ~ N(0,1)
renormalised) and resident on the device.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X dense matrix-core peaks (TFLOP/s)
MFMA_PEAK_TFS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def _cpu_chain(args):
    L, alphar, J2, nsweeps, nmeas, seed = args
    import oracle
    from rbm_pp_amd import lattice as lt
    spec = lt.j1j2_model(L, alphar=alphar, J2=J2)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    orc = oracle.Oracle(spec, p)
    rng = oracle.Rng("mt", seed)
    t0 = time.perf_counter()
    # the reference's adaptive recompute period (fixed_k = 0), as the GPU run
    s = orc.make_sample(rng, True, nsample=nsweeps, nwarmup=0, fixed_k=0)
    t_s = time.perf_counter() - t0
    # calc_phys_quantity_opt on the chain's first samples (Pf, X^-1 recomputed
    # per sample, E_loc, O_k); the S product is timed separately with BLAS
    t0 = time.perf_counter()
    orc.measure(s["x"][:nmeas], shift=0.0, want_s=False)
    t_m = time.perf_counter() - t0
    return nsweeps, t_s, nmeas, t_m


def _host_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(L, alphar, J2, nsweeps, nmeas, procs, ns_step):
    """The oracle's restatement of one SR iteration on the host cores: one
    single-threaded chain per process (the reference's one-rank-per-core
    layout) timed over make_sample and calc_phys_quantity_opt, plus the S
    product and the dense solve with the host BLAS / LAPACK (numpy) on a
    bounded number of rows, scaled to the step's ns_step rows.  Returns
    (end-to-end samples/s, details)."""
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_chain, [(L, alphar, J2, nsweeps, nmeas, 1000 + r) for r in range(procs)])
    wall = time.perf_counter() - t0
    r_samp = procs * nsweeps / max(r[1] for r in res)  # samples/s of the slowest chain x chains
    r_meas = procs * nmeas / max(r[3] for r in res)
    from rbm_pp_amd import lattice as lt
    nv = lt.j1j2_model(L, alphar=alphar, J2=J2).nv
    rows = 20000
    O = np.random.default_rng(0).standard_normal((rows, nv))
    t0 = time.perf_counter()
    S = O.T @ O
    t_s_rows = time.perf_counter() - t0
    A = S / rows + 1e-3 * np.eye(nv)
    t0 = time.perf_counter()
    np.linalg.solve(A, np.ones(nv))
    t_solve = time.perf_counter() - t0
    t_iter = ns_step / r_samp + ns_step / r_meas + t_s_rows * ns_step / rows + t_solve
    det = dict(sampler_samples_per_s=r_samp, measure_samples_per_s=r_meas,
               s_build_s=t_s_rows * ns_step / rows, solve_s=t_solve, wall_s=wall)
    return ns_step / t_iter, det


def synthetic_orbitals(L, nsl):
    """Orbital table of a pair product with nsl orbitals on an L x L lattice
    (the SR step's size only): the diagonal pairs (i, i) share nsl / N
    orbitals, every unordered pair i < j one of the others."""
    N = L * L
    nd = nsl // N
    orb = np.zeros((N, N), np.int64)
    for i in range(N):
        orb[i, i] = i % nd
    iu = np.triu_indices(N, 1)
    idx = nd + np.arange(len(iu[0])) % (nsl - nd)
    orb[iu] = idx
    orb[(iu[1], iu[0])] = idx
    return orb


def sampler_form(projection, N, fused):
    """True when the sampler keeps F^-1 per copy (block form): the normal
    projection (every X = [[0, F], [-F^T, 0]]) up to N = 104 in every sampler
    variant and up to N = 256 without the fused measurement; otherwise X^-1
    in folded skew storage."""
    return projection == "normal" and (N <= 104 or (N <= 256 and not fused))


def sweep_bytes(block, C, N, acc):
    """Algorithmic HBM bytes of one walker sweep (N proposals) at acceptance
    rate acc.  Folded skew storage holds the N^2/2 independent entries of
    each copy's X^-1: every proposal reads them once (8 C N^2/2 bytes) and an
    accepted one reads and writes them once more, 4 C N^3 (1 + 2 acc).  In
    block form the sampler keeps F^-1, (N/2)^2 entries per copy, and applies
    an accepted move's update inside the next proposal's read pass: each
    proposal reads them once and an accepted one writes them once,
    2 C N^3 (1 + acc)."""
    return 2.0 * C * N ** 3 * (1.0 + acc) if block else 4.0 * C * N ** 3 * (1.0 + 2.0 * acc)


def refine_sweeps(name):
    """Refinement sweeps of the SR solve captured in its graph: one.  A
    mixed-precision S is refined against the FP64 operator of the O rows
    until the residual is below 1e-13, checked after every further sweep."""
    return 1


def sr_shape_bench(api, lt, L, alphar, J2, ns, nsl=None, seed=5):
    """The SR step (S build + system solve) on a synthetic measurement of ns
    configurations in the three S precisions: times, FLOP rate of the S build
    against the precision's matrix-core peak, S and step error against FP64."""
    spec = lt.j1j2_model(L, alphar=alphar, J2=J2)
    if nsl is not None:
        spec.orb = synthetic_orbitals(L, nsl)
        spec.n_slater = nsl
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    eng = api.Engine(spec, n_walkers=1, n_samples=1, params=p)
    nv = spec.nv
    eng.set_measured(ns, seed=seed)
    f_s = nv * (nv + 1) * float(ns)
    peaks = {"fp64": MFMA_PEAK_TFS["fp64"], "fp32": MFMA_PEAK_TFS["fp32"], "bf16x3": MFMA_PEAK_TFS["bf16"] / 3.0}
    res = {}
    for name, prec in (("fp64", api.S_FP64), ("fp32", api.S_FP32), ("bf16x3", api.S_BF16X3)):
        eng.set_s_precision(prec)
        best = None
        for _ in range(2):  # the first call also captures the solve's graph
            g, rep = eng.sr_step(p.gamma(), s_factor0=1e-4, delta_tau=0.05, refine_iters=refine_sweeps(name))
            t = eng.timings()
            if best is None or t["sr_total"] < best[1]["sr_total"]:
                best = (g - p.gamma(), t, rep)
        # the same step refined against the assembled mixed-precision matrix
        ga = None
        if name != "fp64":
            ga, _ = eng.sr_step(p.gamma(), s_factor0=1e-4, delta_tau=0.05, refine_iters=1, assembled_refine=True)
            ga = ga - p.gamma()
        S = eng.smat() if nv <= 6000 else None
        res[name] = dict(step=best[0], t=best[1], rep=best[2], S=S, step_assembled=ga)
        log(f"SR step nv={nv} ns={ns} {name}: S build {best[1]['s_build'] * 1e3:.1f} ms, solve "
            f"{best[1]['solve'] * 1e3:.1f} ms, residual {best[2]['residual']:.1e}")
    base = res["fp64"]
    out = {"nv": nv, "configurations": ns, "lattice": f"{L}x{L}", "alphar": alphar, "n_slater": spec.n_slater}
    for name, r in res.items():
        ach = f_s / r["t"]["s_build"] / 1e12
        d = dict(s_build_s=r["t"]["s_build"], solve_s=r["t"]["solve"], sr_step_s=r["t"]["sr_total"],
                 tflops=ach, peak_tflops=peaks[name], mfma_frac=ach / peaks[name],
                 residual=r["rep"]["residual"], cg_iters=r["rep"]["cg_iters"],
                 precond_boost=r["rep"]["precond_boost"],
                 step_rel_err=float(np.linalg.norm(r["step"] - base["step"]) / np.linalg.norm(base["step"])),
                 speedup_vs_fp64=base["t"]["sr_total"] / r["t"]["sr_total"],
                 refine_sweeps=refine_sweeps(name))
        if r["step_assembled"] is not None:
            d["step_rel_err_assembled_refine"] = float(np.linalg.norm(r["step_assembled"] - base["step"]) /
                                                       np.linalg.norm(base["step"]))
        if r["S"] is not None:
            d["s_rel_err"] = float(np.linalg.norm(r["S"] - base["S"]) / np.linalg.norm(base["S"]))
        out[name] = d
    eng.close()
    return out


def run_sr_shapes(api, lt, sizes="both"):
    """The 16x16 workload's SR step (alphar = 16, Nv = 5136, 4x10^5
    configurations) and a synthetic Nv = 19984 (4x10^5 configurations)."""
    out = []
    if sizes in ("both", "5136"):
        out.append(sr_shape_bench(api, lt, 16, 16, 0.55, 400000))
    if sizes in ("both", "19984"):
        out.append(sr_shape_bench(api, lt, 16, 16, 0.55, 400000, nsl=62 * 256))
    return out


def energy_run(args):
    """One full optimisation of the bench workload through the Fortran driver
    (main.f90:300-574 on the C ABI: args.energy_iters SR iterations with S in
    the chosen precision, parameters averaged over the last args.energy_nav,
    then args.energy_nav measurement iterations at fixed parameters), from the
    same seeds for every precision.  Prints one JSON line with the final energy
    per site, its error bar and the per-iteration history."""
    import subprocess
    import tempfile
    from rbm_pp_amd import lattice as lt
    driver = os.path.join(ROOT, "rbm_pp_amd", "fortran", "rbmpp_vmc")
    spec = lt.j1j2_model(args.L, alphar=args.alphar, J2=args.J2)
    nspw = math.ceil(args.samples / args.walkers)
    spec.nsample, spec.nstep, spec.nav = args.walkers * nspw, args.energy_iters, args.energy_nav
    spec.delta_tau, spec.s_factor0, spec.s_factor_scale = 0.05, 1e-4, 1.0
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    rundir = args.energy_dir or tempfile.mkdtemp(prefix="rbmpp_energy_")
    os.makedirs(rundir, exist_ok=True)
    lt.write_run_dir(rundir, spec, p)
    prec = {"fp64": 0, "fp32": 1, "bf16x3": 2}[args.energy_run]
    N = spec.N
    log(f"energy run: {args.L}x{args.L}, S {args.energy_run}, {args.energy_iters} SR iterations, "
        f"Nav {args.energy_nav}, {args.walkers} walkers x {nspw} samples, run dir {rundir}")
    t0 = time.perf_counter()
    proc = subprocess.Popen([driver, str(args.walkers), str(prec), str(args.recompute_k), "0", "0",
                             str(args.energy_seed)], cwd=rundir,
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    it = 0
    sr_res, sr_cg, sr_boost = [], [], []  # per iteration: the SR report of the step
    for line in proc.stdout:
        if line.startswith("Iteration"):
            it = line.split(":")[-1].strip()
        elif "Total Energy" in line or "total energy" in line:
            log(f"iteration {it}: E/N = {float(line.split(':')[-1]) / 4 / N:.6f}")
        elif "SR residual, CG, boost" in line:
            f = line.split(":")[-1].split()
            sr_res.append(float(f[0])); sr_cg.append(int(f[1])); sr_boost.append(float(f[2]))
            if sr_boost[-1] > 0 or sr_res[-1] > 1e-11:
                log(f"iteration {it}: SR residual {sr_res[-1]:.2e}, {sr_cg[-1]} CG iterations, boost {sr_boost[-1]:.2e}")
        elif "failed" in line or "error" in line.lower():
            log(f"driver: {line.rstrip()}")
    rc = proc.wait()
    if rc != 0:
        log(f"driver exit {rc}")
        sys.exit(rc)
    hist = np.loadtxt(os.path.join(rundir, "Energy_vs_Iteration.txt"))
    e_av, e_err = [float(t) for t in open(os.path.join(rundir, "Average_Energy.txt")).readline().split()[:2]]
    print(json.dumps({"energy_run": args.energy_run, "seed": args.energy_seed,
                      "workload": f"j1j2_{args.L}x{args.L}_J2_{args.J2}_alphar{args.alphar}",
                      "iterations": args.energy_iters, "nav": args.energy_nav,
                      "samples_per_iteration": args.walkers * nspw,
                      "energy_per_site": e_av / 4 / N, "error_per_site": e_err / 4 / N,
                      "history_per_site": [float(v) / 4 / N for v in hist[:, 1]],
                      "sr_residual_max": max(sr_res) if sr_res else None,
                      "sr_cg_iters_max": max(sr_cg) if sr_cg else None,
                      "sr_boosted_steps": sum(1 for b in sr_boost if b > 0),
                      "sr_cg_iters": sr_cg, "sr_residual": sr_res,
                      "wall_s": time.perf_counter() - t0}))


def step_compare(args):
    """The mixed-precision SR steps against the FP64 step along one
    optimisation of the bench workload, on the same samples: every iteration
    samples and measures once (main.f90:323-370), forms the SR step from those
    measurements with FP32 and split-BF16 x 3 S products and then with FP64 S
    (SR_pdposv.f90), and continues with the FP64 step.  Reports per iteration
    the mixed steps' relative error ||dg_mixed - dg_fp64|| / ||dg_fp64|| and
    the energy change their difference makes to first order,
    g . (gamma_mixed - gamma_fp64) with g the energy gradient of the same
    samples (SR_pdposv.f90:4-6), per site in the units of E/N.  Prints one
    JSON line."""
    from rbm_pp_amd import api
    from rbm_pp_amd import lattice as lt
    spec = lt.j1j2_model(args.L, alphar=args.alphar, J2=args.J2)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    nspw = math.ceil(args.samples / args.walkers)
    eng = api.Engine(spec, n_walkers=args.walkers, n_samples=nspw, seed=args.energy_seed,
                     recompute_k=args.recompute_k, s_precision=api.S_FP64, params=p, cond_guard=args.cond_guard)
    N = spec.N
    log(f"step comparison: {args.L}x{args.L}, {args.step_compare} iterations, {args.walkers} walkers x {nspw} "
        f"samples, nv={spec.nv}")
    modes = (("fp32", api.S_FP32), ("bf16x3", api.S_BF16X3), ("fp64", api.S_FP64))
    eng.make_sample(True, shift=0.0)
    wf_list, shift, gam = [0.0] * 10, 0.0, eng.gamma()
    hist = []
    t0 = time.perf_counter()
    for it in range(args.step_compare):
        eng.make_sample(False, shift=shift)
        r = eng.calc_phys_quantity_opt(shift=shift)
        grad = r["gvec"] - 2.0 * r["Etot"] * r["ovec"]
        res = {}
        for name, prec in modes:  # FP64 last: its step is the one the engine keeps
            eng.set_s_precision(prec)
            g_new, rep = eng.sr_step(gam, s_factor0=1e-4, delta_tau=0.05, refine_iters=refine_sweeps(name))
            res[name] = (g_new, rep, eng.timings()["sr_total"])
        g64 = res["fp64"][0]
        d64 = g64 - gam
        row = {"E_per_site": r["Etot"] / 4 / N, "sr_fp64_s": res["fp64"][2]}
        for name in ("fp32", "bf16x3"):
            dm = res[name][0] - g64
            row[name] = {"step_rel_err": float(np.linalg.norm(dm) / np.linalg.norm(d64)),
                         "dE_per_site": float(grad @ dm) / 4 / N, "cg_iters": res[name][1]["cg_iters"],
                         "boost": res[name][1]["precond_boost"], "sr_s": res[name][2]}
        hist.append(row)
        gam = g64
        wf_list.pop(0)
        wf_list.append(r["wf_av"])
        shift = sum(wf_list) / 10.0 * 1e-3 / N / N
        log(f"iteration {it}: E/N = {row['E_per_site']:.6f}, step error fp32 {row['fp32']['step_rel_err']:.1e} "
            f"bf16x3 {row['bf16x3']['step_rel_err']:.1e}, dE/N bf16x3 {row['bf16x3']['dE_per_site']:.1e}")
    out = {"step_compare": f"j1j2_{args.L}x{args.L}_J2_{args.J2}_alphar{args.alphar}", "nv": spec.nv,
           "iterations": args.step_compare, "samples_per_iteration": args.walkers * nspw,
           "E_per_site_first": hist[0]["E_per_site"], "E_per_site_last": hist[-1]["E_per_site"],
           "wall_s": time.perf_counter() - t0}
    for name in ("fp32", "bf16x3"):
        out[name] = {"step_rel_err_max": max(h[name]["step_rel_err"] for h in hist),
                     "abs_dE_per_site_max": max(abs(h[name]["dE_per_site"]) for h in hist),
                     "cg_iters_max": max(h[name]["cg_iters"] for h in hist),
                     "boosted_steps": sum(1 for h in hist if h[name]["boost"] > 0),
                     "sr_s_mean": float(np.mean([h[name]["sr_s"] for h in hist]))}
    out["fp64_sr_s_mean"] = float(np.mean([h["sr_fp64_s"] for h in hist]))
    out["history"] = hist
    print(json.dumps(out))


def batch_run(args):
    """The uneven multi-lattice workload: L = 6, 8, 10, 12, 14 at J2/J1 = 0.4,
    0.5, 0.55 (15 models), RBM alphar = 2, 2x2-sublattice pair product, C4v +
    parity, args.batch_walkers walkers apportioned by N^2.  Each iteration:
    one sampler launch over all walkers (rbmpp_batch_make_sample), then per
    model the measurement and an SR step with split-BF16 x 3 S.  The same
    sweeps launched model by model are timed once for comparison.  Prints one
    JSON line."""
    from rbm_pp_amd import api
    from rbm_pp_amd import lattice as lt
    models = [(L, J2) for L in (6, 8, 10, 12, 14) for J2 in (0.4, 0.5, 0.55)]
    n2 = np.array([float(L) ** 4 for L, _ in models])
    walkers = np.maximum(1, np.round(args.batch_walkers * n2 / n2.sum())).astype(int)
    engines = []
    for (L, J2), W in zip(models, walkers):
        spec = lt.j1j2_model(L, alphar=2, J2=J2)
        p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
        engines.append(api.Engine(spec, params=p, n_walkers=int(W), n_samples=args.batch_samples, seed=L,
                                  n_warmup=args.batch_warmup, n_warmup_first=args.batch_warmup,
                                  s_precision=api.S_BF16X3, cond_guard=args.cond_guard))
    batch = api.Batch(engines)
    nmod = len(engines)
    log(f"batch: {nmod} models, walkers {walkers.tolist()} (total {int(walkers.sum())}), "
        f"{args.batch_samples} samples + {args.batch_warmup} warm-up sweeps per walker and iteration")
    shifts = np.zeros(nmod)
    batch.make_sample([1] * nmod, shifts)  # first call: initial configurations
    t_launch, t_meas, t_sr, t_wall, E = [], [], [], [], None
    for it in range(args.batch_run):
        t0 = time.perf_counter()
        batch.make_sample([0] * nmod, shifts)
        t_launch.append(batch.sample_seconds())
        tm = ts = 0.0
        E = []
        for k, e in enumerate(engines):
            a = time.perf_counter()
            r = e.calc_phys_quantity_opt(shift=shifts[k])
            b = time.perf_counter()
            e.sr_step(e.gamma(), s_factor0=1e-4, delta_tau=0.05, refine_iters=refine_sweeps("bf16x3"))
            tm += b - a
            ts += time.perf_counter() - b
            E.append(r["Etot"] / 4 / e.N)
            shifts[k] = r["wf_av"] * 1e-3 / e.N / e.N
        t_meas.append(tm)
        t_sr.append(ts)
        t_wall.append(time.perf_counter() - t0)
        log(f"iteration {it}: {t_wall[-1]:.2f} s (sampler launch {t_launch[-1]:.2f} s, measurements {tm:.2f} s, "
            f"SR steps {ts:.2f} s)")
    # the same sweeps model by model
    t_sep = 0.0
    for k, e in enumerate(engines):
        e.make_sample(False, shift=shifts[k])
        t_sep += e.timings()["sample"]
    log(f"model-by-model launches: {t_sep:.2f} s against {t_launch[-1]:.2f} s in one launch")
    samples = float(walkers.sum()) * args.batch_samples
    print(json.dumps({"batch_run": "uneven multi-lattice", "models": [f"{L}x{L} J2={J2}" for L, J2 in models],
                      "walkers": walkers.tolist(), "samples_per_iteration": samples,
                      "iterations": args.batch_run, "seconds_per_iteration": float(np.mean(t_wall)),
                      "samples_per_s": samples / float(np.mean(t_wall)),
                      "sampler_launch_s": float(np.mean(t_launch)), "model_by_model_launches_s": t_sep,
                      "measurements_s": float(np.mean(t_meas)), "sr_steps_s": float(np.mean(t_sr)),
                      "energy_per_site_last": E}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=10)
    ap.add_argument("--alphar", type=int, default=4)
    ap.add_argument("--J2", type=float, default=0.5)
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--samples", type=int, default=200000, help="stored samples per SR step")
    ap.add_argument("--therm", type=int, default=200,
                    help="thermalisation sweeps, run as make_sample launches of the step's shape")
    ap.add_argument("--recompute-k", type=int, default=0,
                    help="recompute Pf and X^-1 every K accepts; 0: the reference's adaptive period "
                         "(make_sample.f90:74-128, at most 100)")
    ap.add_argument("--cpu-sweeps", type=int, default=600)
    ap.add_argument("--cpu-procs", type=int, default=0)
    ap.add_argument("--cpu-meas", type=int, default=60, help="samples measured per CPU chain")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--s-precision", choices=["fp64", "fp32", "bf16x3"], default="fp32",
                    help="S products of the timed SR steps (centring, shift and solve stay FP64)")
    ap.add_argument("--fused-measure", action="store_true",
                    help="measure inside the sampler with the fast-updated X^-1 instead of "
                         "recomputing Pf and X^-1 per sample as the reference does")
    ap.add_argument("--cond-guard", type=float, default=1e4,
                    help="recompute after an accept whose Pf ratio of a copy leaves (1/g, g); 0: off")
    ap.add_argument("--projection", choices=["normal", "normal_exact", "transverse"], default="normal",
                    help="spin-projection weights: normal (cos(pi/2) = 0, block-form sampler), normal_exact "
                         "(the reference's cos(pi/2) = 6.1e-17, skew-form sampler), transverse")
    ap.add_argument("--no-correlation", action="store_true",
                    help="skip the correlation-run measurement on the last step's samples")
    ap.add_argument("--no-sr-shapes", action="store_true",
                    help="skip the SR steps at the 16x16 sizes (Nv = 5136 and a synthetic ~2e4)")
    ap.add_argument("--sr-only", action="store_true",
                    help="run only the SR steps at the large sizes and print them")
    ap.add_argument("--sr-sizes", choices=["both", "5136", "19984"], default="both",
                    help="which of the large SR shapes --sr-only runs")
    ap.add_argument("--energy-run", choices=["fp64", "fp32", "bf16x3"],
                    help="one full optimisation through the Fortran driver with this S precision; prints "
                         "the final energy per site and its error bar")
    ap.add_argument("--energy-iters", type=int, default=250)
    ap.add_argument("--energy-nav", type=int, default=20)
    ap.add_argument("--energy-dir", default=None, help="run directory of --energy-run")
    ap.add_argument("--energy-seed", type=int, default=1, help="key of the walkers' Philox streams in --energy-run")
    ap.add_argument("--step-compare", type=int, default=0, metavar="ITERS",
                    help="an optimisation of ITERS iterations continuing with the FP64 step, each iteration's "
                         "FP32 and split-BF16 x 3 steps compared with it on the same samples")
    ap.add_argument("--batch-run", type=int, default=0, metavar="ITERS",
                    help="the uneven multi-lattice workload (15 models in one sampler launch) for ITERS iterations")
    ap.add_argument("--batch-walkers", type=int, default=16384)
    ap.add_argument("--batch-samples", type=int, default=4)
    ap.add_argument("--batch-warmup", type=int, default=4)
    ap.add_argument("--sampler-only", action="store_true",
                    help="profiling aid: thermalise, run one make_sample launch and exit")
    ap.add_argument("--with-measure", action="store_true",
                    help="with --sampler-only: also time the measurement of the launch's samples")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="profiling aid: three device-to-device copies of MIB MiB and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        # a plain copy of known size, to scale the HBM byte counters
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}))
        return

    if args.energy_run:
        energy_run(args)
        return
    if args.step_compare:
        step_compare(args)
        return
    if args.batch_run:
        batch_run(args)
        return

    if args.sr_only:
        from rbm_pp_amd import api
        from rbm_pp_amd import lattice as lt
        print(json.dumps({"sr_shapes": run_sr_shapes(api, lt, args.sr_sizes)}))
        return

    # CPU baseline first: worker processes are forked before the GPU is touched
    cpu = None
    if not args.no_cpu and not args.sampler_only:
        # the GPU box allots 16 host cores to its one GPU
        procs = args.cpu_procs or min(16, os.cpu_count() or 1)
        model, ncpu = _host_info()
        ns_step = args.walkers * math.ceil(args.samples / args.walkers)
        log(f"cpu baseline: {procs} chains x {args.cpu_sweeps} sweeps + {args.cpu_meas} measured samples each "
            f"(oracle), host {model}, {ncpu} logical CPUs")
        rate, det = cpu_baseline(args.L, args.alphar, args.J2, args.cpu_sweeps, args.cpu_meas, procs, ns_step)
        cal = None
        try:
            from oracle import calibrate
            from rbm_pp_amd import lattice as lt0
            spec0 = lt0.j1j2_model(args.L, alphar=args.alphar, J2=args.J2)
            cal = calibrate.fast_update_times(spec0, lt0.random_params(spec0, seed_w=1, seed_f=2, sigma_w=0.01))
        except OSError:
            cal = None
        cpu = {"value": rate, "unit": "samples/s", "cores": procs, "kind": "port",
               "sample": f"{procs} chains (one per core, OMP_NUM_THREADS=1 each) x {args.cpu_sweeps} stored sweeps "
                         f"(N={args.L * args.L} proposals each, adaptive recompute period, make_sample.f90:74-128) "
                         f"and calc_phys_quantity_opt on {args.cpu_meas} samples per chain; S = O^T O on 20000 "
                         f"rows and the dense solve with numpy BLAS/LAPACK, scaled to {ns_step} samples per SR "
                         f"iteration; {det['wall_s']:.1f} s wall",
               "host_cpu": model, "host_logical_cpus": ncpu,
               "sampler_samples_per_s": det["sampler_samples_per_s"],
               "measure_samples_per_s": det["measure_samples_per_s"],
               "s_build_s": det["s_build_s"], "solve_s": det["solve_s"]}
        if cal:
            # seconds per call of pfupdate_two.f90's two routines, the
            # restatement over the compiled reference on identical inputs
            cpu["calibration"] = {"pfm_two_port_over_reference": cal["port_ratio_s"] / cal["ref_ratio_s"],
                                  "invm_two_port_over_reference": cal["port_update_s"] / cal["ref_update_s"],
                                  "pfm_two_reference_s": cal["ref_ratio_s"], "invm_two_reference_s": cal["ref_update_s"]}
        log(f"cpu baseline: {rate:.1f} samples/s end to end (sampler {det['sampler_samples_per_s']:.1f}/s, "
            f"measurement {det['measure_samples_per_s']:.1f}/s)")

    from rbm_pp_amd import api
    S_MODES = {"fp64": api.S_FP64, "fp32": api.S_FP32, "bf16x3": api.S_BF16X3}
    from rbm_pp_amd import lattice as lt

    spec = lt.j1j2_model(args.L, alphar=args.alphar, J2=args.J2)
    p = lt.random_params(spec, seed_w=1, seed_f=2, sigma_w=0.01)
    nspw = math.ceil(args.samples / args.walkers)
    eng = api.Engine(spec, n_walkers=args.walkers, n_samples=nspw, seed=1, n_warmup=-1,
                     n_warmup_first=-1, recompute_k=args.recompute_k,
                     s_precision=S_MODES[args.s_precision], params=p, fused_measure=args.fused_measure,
                     cond_guard=args.cond_guard)
    eng.set_projection({"normal": api.PROJ_NORMAL, "normal_exact": api.PROJ_NORMAL_EXACT,
                        "transverse": api.PROJ_TRANSVERSE}[args.projection])
    N, Cc = spec.N, spec.C
    log(f"model L={args.L} N={N} copies={Cc} nv={spec.nv} walkers={args.walkers} samples/walker={nspw}")
    # every k_sweep launch has the shape of a step's: nwarm + 1 + nspw sweeps
    nwarm = max(30, nspw // 100)
    sweeps_per_launch_walker = nwarm + 1 + nspw
    t0 = time.perf_counter()
    eng.make_sample(True, shift=0.0)
    done = sweeps_per_launch_walker
    while done < args.therm:
        eng.make_sample(False, shift=0.0)
        done += sweeps_per_launch_walker
    log(f"thermalisation ({done} sweeps/walker): {time.perf_counter() - t0:.1f} s")
    if args.sampler_only:
        nacc = eng.make_sample(False, shift=0.0)
        t = eng.timings()
        acc = nacc / (args.walkers * nspw * N)
        block = sampler_form(args.projection, N, args.fused_measure)
        b_unit = sweep_bytes(block, Cc, N, acc)
        gbs = b_unit * args.walkers * sweeps_per_launch_walker / t["sample"] / 1e9
        out = {"kernel": "k_sweep",
               "workload": f"j1j2_{args.L}x{args.L}_J2_{args.J2}_alphar{args.alphar}_c4v_parity",
               "walkers": args.walkers, "sweeps_per_walker": sweeps_per_launch_walker,
               "launch_s": t["sample"], "acceptance": acc,
               "samples_per_s": args.walkers * nspw / t["sample"],
               "layout": "F^-1 per copy (block form)" if block else "folded skew X^-1 per copy",
               "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": gbs / HBM_PEAK_GBS, "bytes_per_walker_sweep": b_unit}}
        if args.with_measure:
            # calc_phys_quantity_opt on the launch's samples (E_loc, O_k)
            eng.calc_phys_quantity_opt(shift=0.0)
            tm = eng.timings()["measure"]
            out["measure_s"] = tm
            out["measure_samples_per_s"] = args.walkers * nspw / tm
        print(json.dumps(out))
        return

    wf_list = [0.0] * 10
    shift = 0.0
    gam = eng.gamma()
    gam_prev = gam.copy()

    last_grad = None  # the energy gradient of the last measurement (SR_pdposv.f90:4-6)

    def one_step(rec):
        nonlocal shift, gam, gam_prev, last_grad
        t0 = time.perf_counter()
        nacc = eng.make_sample(False, shift=shift)
        tt = eng.timings()
        t1 = time.perf_counter()
        r = eng.calc_phys_quantity_opt(shift=shift)
        t2 = time.perf_counter()
        last_grad = r["gvec"] - 2.0 * r["Etot"] * r["ovec"]
        gam_prev = gam.copy()
        gam, rep = eng.sr_step(gam, s_factor0=1e-4, delta_tau=0.05, refine_iters=refine_sweeps(args.s_precision))
        t3 = time.perf_counter()
        tt2 = eng.timings()
        # main.f90:441-451: reweighting shift from the running mean of wf_av
        wf_list.pop(0); wf_list.append(r["wf_av"])
        shift = sum(wf_list) / 10.0 * 1e-3 / N / N
        if rec is not None:
            rec.append(dict(wall=t3 - t0, sample=t1 - t0, measure=t2 - t1, sr=t3 - t2,
                            sweep_kernel=tt["sample"], s_build=tt2["s_build"], solve=tt2["solve"],
                            recomputes=tt["recompute_count"],
                            nacc=nacc, E=r["Etot"], lam=rep["lambda_"]))
        return nacc, r

    for i in range(args.warmup):
        nacc, r = one_step(None)
        log(f"warmup step {i}: E/N={r['Etot'] / 4 / N:.6f}")
    ns_step = args.walkers * nspw
    recs = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        one_step(recs)
        x = recs[-1]
        log(f"step {i}: {x['wall']:.2f} s (sample {x['sample']:.2f}, measure {x['measure']:.2f}, "
            f"sr {x['sr']:.3f}) E/N={x['E'] / 4 / N:.6f} accept={x['nacc'] / ns_step / N:.3f} "
            f"recomputes/walker={x['recomputes'] / args.walkers:.1f}")
    t_total = time.perf_counter() - t_start

    # The last step's samples through the three S precisions: S-build time and
    # error, and the SR step from the same parameters against the FP64 one.
    # Algorithmic flops of the S build: one triangle, nv (nv + 1) / 2 entries
    # of rows-long dot products (2 flops per term).
    fp64 = None
    if not args.no_fp64_compare:
        rows = ns_step if not spec.alphac else 2 * ns_step
        f_s = spec.nv * (spec.nv + 1) * float(rows)
        peaks = {"fp64": MFMA_PEAK_TFS["fp64"], "fp32": MFMA_PEAK_TFS["fp32"],
                 "bf16x3": MFMA_PEAK_TFS["bf16"] / 3.0}
        res = {}
        for name, prec in (("fp64", api.S_FP64), ("fp32", api.S_FP32), ("bf16x3", api.S_BF16X3)):
            eng.set_s_precision(prec)
            S = eng.smat()
            t_s = eng.timings()["s_build"]
            g_new, _ = eng.sr_step(gam_prev, s_factor0=1e-4, delta_tau=0.05, refine_iters=refine_sweeps(name))
            t_sr = eng.timings()["sr_total"]
            g_asm = None
            if name != "fp64":
                g_asm, _ = eng.sr_step(gam_prev, s_factor0=1e-4, delta_tau=0.05, refine_iters=1,
                                       assembled_refine=True)
            res[name] = dict(S=S, t_s=t_s, t_sr=t_sr, step=g_new - gam_prev,
                             step_asm=None if g_asm is None else g_asm - gam_prev)
        base = res["fp64"]
        fp64 = {}
        for name, r in res.items():
            ach = f_s / r["t_s"] / 1e12
            fp64[name] = dict(s_build_s=r["t_s"], sr_step_s=r["t_sr"], tflops=ach, peak_tflops=peaks[name],
                              mfma_frac=ach / peaks[name],
                              s_rel_err=float(np.linalg.norm(r["S"] - base["S"]) / np.linalg.norm(base["S"])),
                              step_rel_err=float(np.linalg.norm(r["step"] - base["step"]) /
                                                 np.linalg.norm(base["step"])),
                              # the energy change per site that the step's
                              # difference from the FP64 step makes, to first
                              # order, on the same samples
                              dE_per_site=float(last_grad @ (r["step"] - base["step"])) / 4 / N,
                              refine_sweeps=refine_sweeps(name))
            if r["step_asm"] is not None:
                fp64[name]["step_rel_err_assembled_refine"] = float(
                    np.linalg.norm(r["step_asm"] - base["step"]) / np.linalg.norm(base["step"]))
            log(f"S build {name}: {r['t_s'] * 1e3:.3f} ms ({ach:.1f} TFLOP/s, {100 * ach / peaks[name]:.1f}% of "
                f"{peaks[name]:.0f}), SR step {r['t_sr'] * 1e3:.2f} ms, S err {fp64[name]['s_rel_err']:.2e}, "
                f"step err {fp64[name]['step_rel_err']:.2e}, dE/N {fp64[name]['dE_per_site']:.1e}")
        eng.set_s_precision(S_MODES[args.s_precision])

    # The measurement of the correlation run (Calc_Correlation: Chi_sp over
    # all N (N + 1) / 2 pairs, Chi_vbs, wf_sgn) on the last step's samples.
    corr = None
    if not args.no_correlation:
        t0 = time.perf_counter()
        c = eng.measure_correlation(shift=shift)
        wall = time.perf_counter() - t0
        t_dev = eng.timings()["measure"]
        corr = {"samples": ns_step, "measure_s": t_dev, "wall_s": wall, "samples_per_s": ns_step / t_dev,
                "exchange_pairs_per_sample": (N // 2) ** 2, "wf_sgn_abs": abs(c["wf_sgn"]),
                "energy_per_site": c["Etot"] / 4 / N}
        log(f"correlation measurement: {t_dev:.2f} s for {ns_step} samples ({ns_step / t_dev:.0f} samples/s)")

    sr_shapes = None
    if not args.no_sr_shapes:
        sr_shapes = run_sr_shapes(api, lt)

    samples = ns_step * args.steps
    acc = sum(r["nacc"] for r in recs) / samples / N
    sweeps_per_launch = args.walkers * sweeps_per_launch_walker
    t_sweep = sum(r["sweep_kernel"] for r in recs) / len(recs)
    # algorithmic HBM bytes per sweep of one walker (sweep_bytes)
    block = sampler_form(args.projection, N, args.fused_measure)
    b_unit = sweep_bytes(block, Cc, N, acc)
    achieved = b_unit * sweeps_per_launch / t_sweep / 1e9
    # measured HBM bytes per walker sweep of this kernel and workload
    # (rocprofv3 FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a
    # copy of known size), when the committed counter summary covers it
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"k_sweep_{'block' if block else 'skew'}_j1j2_{args.L}x{args.L}_pmc.md")
    if os.path.exists(pmc):
        for line in open(pmc):
            if line.startswith("traffic_bytes_per_walker_sweep:"):
                traffic = float(line.split(":")[1]) * sweeps_per_launch
    out = {
        "metric": "MC samples/s + SR-step s, 10x10 J1-J2; mixed-prec vs FP64 at equal energy",
        "value": samples / t_total,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"j1j2_{args.L}x{args.L}_J2_{args.J2}_alphar{args.alphar}_c4v_parity",
                   "walkers": args.walkers, "samples_per_sr_step": ns_step,
                   "sweeps_per_walker_per_step": sweeps_per_launch_walker,
                   "recompute_k": args.recompute_k,
                   "s_precision": f"{args.s_precision} products on MFMA, fp64 centring/shift/solve",
                   "measurement": "fused into sampler" if args.fused_measure else "recompute per sample",
                   "projection": args.projection,
                   "sampler_layout": "F^-1 per copy (block form)" if block else "folded skew X^-1 per copy",
                   "cond_guard": args.cond_guard,
                   "parallelism": "single_gpu"},
        "sampler_samples_per_s": samples / sum(r["sample"] for r in recs),
        "sr_step_s": sum(r["sr"] for r in recs) / len(recs),
        "measure_s": sum(r["measure"] for r in recs) / len(recs),
        "sample_s": sum(r["sample"] for r in recs) / len(recs),
        "acceptance": acc,
        "energy_per_site": recs[-1]["E"] / 4 / N,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sweep", "launch_s": t_sweep,
                     "bytes_per_walker_sweep": b_unit},
        "cpu_baseline": cpu,
    }
    if fp64:
        out["s_build"] = fp64
    if corr:
        out["correlation"] = corr
    if sr_shapes:
        out["sr_shapes"] = sr_shapes
    print(json.dumps(out))


if __name__ == "__main__":
    main()
