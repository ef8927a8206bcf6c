"""Calibration of the C restatement against the reference's own compiled
routines on identical inputs: calculate_new_PfM_two (every proposal) and
calculate_new_InvM_two (every accepted proposal) of pfupdate_two.f90, which
carry the sampler's cost, from oracle/_ref/libref.so (built from the
reference sources) and from the restatement (liboracle.so), single-threaded.

TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg to state how
the CPU baseline (the restatement) compares with the reference itself."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import Oracle, lib
from . import refharness as RH


def _raw_invm(orc: Oracle) -> np.ndarray:
    """InvM of every copy in the reference's column-major layout."""
    N, Cc = orc.N, orc.C
    bufs = [np.zeros(N, np.int32), np.zeros(2 * max(orc.mc, 1)), np.zeros(max(orc.mr, 1)),
            np.zeros(N, np.int32), np.zeros(N, np.int32), np.zeros(orc.P, np.int32), np.zeros(Cc),
            np.zeros(Cc * N * N), np.zeros(2 * Cc), np.zeros(2 * (Cc + 1))]
    lib().om_state_get(orc.h, orc.st, *[b.ctypes.data_as(C.c_void_p) for b in bufs])
    return bufs[7]


def fast_update_times(spec, params, proposals: int = 200, accepts: int = 40, seed: int = 3) -> dict | None:
    """Seconds per call of the two routines in the restatement and in the
    reference on the same configurations, exchanges and InvM; None when the
    reference routines are not built here."""
    if not RH.available(0):
        return None
    ref = RH.Ref(spec, params)
    orc = Oracle(spec, params)
    L = lib()
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.om_pfm_two.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, dp, dp, dp]
    L.om_invm_two.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, dp]
    rng = np.random.default_rng(seed)
    N, Cc = spec.N, spec.C
    x = np.array([1] * (N // 2) + [-1] * (N // 2), np.int32)
    rng.shuffle(x)
    st = orc.eval(x)
    invm = _raw_invm(orc)
    pf = np.ascontiguousarray(st["pf"])
    idx, cfg = st["ele_idx"].astype(np.int32), st["ele_cfg"]
    moves = []
    for _ in range(proposals):
        i1 = int(rng.integers(N))
        i2 = int(rng.choice(np.nonzero(x != x[i1])[0]))
        msa, msb = int(cfg[i1]), int(cfg[i2])
        tmp = idx.copy()
        tmp[msa], tmp[msb] = i2, i1
        moves.append((msa, msb, np.ascontiguousarray(tmp)))
    out_p, out_r = np.zeros(Cc), np.zeros(Cc)
    # identical results first (the timing compares the same computation)
    msa, msb, tmp = moves[0]
    L.om_pfm_two(orc.h, msa, msb, tmp, invm, pf, out_p)
    ref.L.ref_pfm_two(msa, msb, tmp, invm, pf, out_r)
    agree = float(np.abs(out_p - out_r).max() / np.abs(out_r).max())
    t0 = time.perf_counter()
    for msa, msb, tmp in moves:
        L.om_pfm_two(orc.h, msa, msb, tmp, invm, pf, out_p)
    t_port_ratio = (time.perf_counter() - t0) / proposals
    t0 = time.perf_counter()
    for msa, msb, tmp in moves:
        ref.L.ref_pfm_two(msa, msb, tmp, invm, pf, out_r)
    t_ref_ratio = (time.perf_counter() - t0) / proposals
    # the update on a fresh copy of InvM each time (the copy is timed alike)
    t_port_upd = t_ref_upd = 0.0
    for msa, msb, tmp in moves[:accepts]:
        w = invm.copy()
        t0 = time.perf_counter()
        L.om_invm_two(orc.h, msa, msb, tmp, w)
        t_port_upd += time.perf_counter() - t0
        w = invm.copy()
        t0 = time.perf_counter()
        ref.L.ref_invm_two(msa, msb, tmp, w)
        t_ref_upd += time.perf_counter() - t0
    return dict(port_ratio_s=t_port_ratio, ref_ratio_s=t_ref_ratio,
                port_update_s=t_port_upd / accepts, ref_update_s=t_ref_upd / accepts,
                ratio_agreement=agree)
