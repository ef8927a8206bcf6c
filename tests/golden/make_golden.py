"""Writes the golden fixtures in tests/golden/ from the reference's own
routines (oracle/_ref/libref.so, built by oracle/Makefile from the reference
sources).  Run in a container that holds the reference sources:

    make -C oracle ref && python tests/golden/make_golden.py [name ...]

Inputs are generated here from fixed seeds.  Pf and InvM are inputs to the
reference routines and come from the CPU oracle, because the reference's own
calculate_MAll needs the prebuilt PFAPACK archive and is not built.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from oracle import refharness as RH  # noqa: E402
from rbm_pp_amd import lattice as lt  # noqa: E402

CASES = {
    "ref_4x4_a1c0": dict(L=4, alphar=1, alphac=0, J2=0.5, nconf=8, seed=11),
    "ref_4x4_a2c1": dict(L=4, alphar=2, alphac=1, J2=0.5, nconf=8, seed=12),
    "ref_6x6_a2c0": dict(L=6, alphar=2, alphac=0, J2=0.5, nconf=4, seed=13),
    # spin weights of transverse_code/spin_projection.f90
    "ref_4x4_a2c1_tc": dict(L=4, alphar=2, alphac=1, J2=0.5, nconf=8, seed=14, projection=1),
}


def main(names=None):
    for name, c in CASES.items():
        if names and name not in names:
            continue
        proj = c.get("projection", 0)
        spec = lt.j1j2_model(c["L"], alphar=c["alphar"], alphac=c["alphac"], J2=c["J2"])
        p = lt.random_params(spec, seed_w=c["seed"], seed_f=c["seed"] + 100, sigma_w=0.05)
        orc = O.Oracle(spec, p)
        orc.set_projection(proj)
        ref = RH.Ref(spec, p, projection=proj)
        rng = np.random.default_rng(c["seed"])
        N = spec.N
        out = dict(L=c["L"], alphar=c["alphar"], alphac=c["alphac"], J2=c["J2"], projection=proj,
                   Wc=p.Wc, bc=p.bc, Wr=p.Wr, br=p.br, f=p.f,
                   slater_elm_q0=ref.slater_elm()[0], sp=ref.sp())
        keys = ["x", "thetac", "thetar", "ele_idx", "ele_cfg", "sgn", "pf_in", "psi",
                "eloc", "opp", "ex_i1", "ex_i2", "pf_two", "psi_two"]
        rec = {k: [] for k in keys}
        invm_upd = []
        for k in range(c["nconf"]):
            x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
            rng.shuffle(x)
            st = orc.eval(x)
            thc, thr = ref.theta(x)
            idx, cfg = ref.x_to_ele(x)
            sg = ref.sign_pp(idx)
            pwo, psi = ref.amplitude(thc, thr, st["pf"], sg)
            st_ref = dict(st, thetac=thc, thetar=thr, ele_idx=idx, ele_cfg=cfg, sgn=sg, psi=psi,
                          psi_wo=pwo)
            e = ref.eloc(st_ref)
            opp = (ref.deriv_pf(st_ref) * pwo[:, None]).sum(0)
            i1 = int(rng.integers(N))
            i2 = int(rng.choice(np.nonzero(x != x[i1])[0]))
            tmp = idx.copy(); msa, msb = cfg[i1], cfg[i2]; tmp[msa] = i2; tmp[msb] = i1
            pft = ref.pfm_two(msa, msb, tmp, st["invm"], st["pf"])
            thc2, thr2 = ref.update_theta(x, i1, i2, thc, thr)
            sg2 = -sg * (-1 if lt.sublattice_b(spec.L)[i1] else 1) * (-1 if lt.sublattice_b(spec.L)[i2] else 1)
            _, psi2 = ref.amplitude(thc2, thr2, pft, sg2)
            if c["L"] == 4 and k < 2:
                invm_upd.append(ref.invm_two(msa, msb, tmp, st["invm"]))
            for kk, v in zip(keys, [x, thc, thr, idx, cfg, sg, st["pf"], psi, e, opp, i1, i2, pft, psi2]):
                rec[kk].append(v)
        for kk in keys:
            out[kk] = np.array(rec[kk])
        if invm_upd:
            out["invm_two"] = np.array(invm_upd)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name)
    if not names or "ref_4x4_a2c1_cc" in names:
        correlation_case("ref_4x4_a2c1_cc", L=4, alphar=2, alphac=1, nconf=8, seed=15)
    if names and "ref_mt19937" not in names:
        return
    # mt.f stream
    spec = lt.j1j2_model(4, alphar=1, J2=0.5)
    ref = RH.Ref(spec, lt.random_params(spec))
    np.savez_compressed(os.path.join(HERE, "ref_mt19937.npz"),
                        seed1=ref.grnd(1, 2000), seed4357=ref.grnd(4357, 2000),
                        seed123456789=ref.grnd(123456789, 2000))
    print("wrote ref_mt19937")


def correlation_case(name, L, alphar, alphac, nconf, seed):
    """Per-configuration SdotS and Chi_loc_x of the Calc_Correlation tree
    (its calc_local_energy_correlation and calc_local_correlation)."""
    spec = lt.j1j2_model(L, alphar=alphar, alphac=alphac, J2=0.5)
    p = lt.random_params(spec, seed_w=seed, seed_f=seed + 100, sigma_w=0.05)
    orc = O.Oracle(spec, p)
    ref = RH.Ref(spec, p, tree="cc")
    rng = np.random.default_rng(seed)
    N = spec.N
    xs, el, sd, cl = [], [], [], []
    for _ in range(nconf):
        x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
        rng.shuffle(x)
        st = orc.eval(x)
        e, s4 = ref.eloc_sdots(st)
        xs.append(x); el.append(e); sd.append(s4); cl.append(ref.local_correlation(st))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), L=L, alphar=alphar, alphac=alphac, J2=0.5,
                        Wc=p.Wc, bc=p.bc, Wr=p.Wr, br=p.br, f=p.f, x=np.array(xs), eloc=np.array(el),
                        sdots=np.array(sd), chi_loc=np.array(cl))
    print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1:])
