"""ctypes front end of the CPU oracle (liboracle.so, built from rbmpp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline, never by the
product package rbm_pp_amd.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "rbmpp_oracle.c")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-shared", "-fPIC",
                               "-o", LIB, SRC, "-lm"])
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.om_rng_mt.restype = P; L.om_rng_mt.argtypes = [C.c_int32]
        L.om_rng_philox.restype = P
        L.om_rng_philox.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        L.om_rng_free.argtypes = [P]
        L.om_rng_u01.restype = C.c_double; L.om_rng_u01.argtypes = [P]
        L.om_rng_fill.argtypes = [P, C.c_int, dp]
        L.om_rng_counter.restype = C.c_uint64; L.om_rng_counter.argtypes = [P]
        u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        L.om_philox4x32_10.argtypes = [u32, u32, u32]
        L.om_create.restype = P
        L.om_create.argtypes = [C.c_int] * 5 + [C.c_double, C.c_double, C.c_int, ip, dp, ip]
        L.om_free.argtypes = [P]
        L.om_nv.restype = C.c_int; L.om_nv.argtypes = [P]
        L.om_nv_rbm.restype = C.c_int; L.om_nv_rbm.argtypes = [P]
        L.om_set_params.argtypes = [P, dp, dp, dp, dp, dp]
        L.om_set_gamma.argtypes = [P, dp]
        L.om_set_projection.argtypes = [P, C.c_int]
        L.om_get_gamma.argtypes = [P, dp]
        L.om_slater_elm.restype = C.POINTER(C.c_double); L.om_slater_elm.argtypes = [P]
        L.om_sp_weights.argtypes = [P, dp]
        L.om_pfaffian.restype = C.c_double; L.om_pfaffian.argtypes = [C.c_int, dp]
        L.om_inverse.argtypes = [C.c_int, dp, dp]
        L.om_state_create.restype = P; L.om_state_create.argtypes = [P]
        L.om_state_free.argtypes = [P]
        L.om_eval.argtypes = [P, ip, P, P]
        L.om_state_get.argtypes = [P, P] + [C.c_void_p] * 10
        L.om_calc_theta.argtypes = [P, ip, dp, dp]
        L.om_sign_pp.argtypes = [P, ip, ip]
        L.om_ratio_exchange.argtypes = [P, P, C.c_int, C.c_int, dp, dp, dp]
        L.om_accept_exchange.argtypes = [P, P, C.c_int, C.c_int, P]
        L.om_eloc.restype = C.c_double; L.om_eloc.argtypes = [P, P, dp, P]
        L.om_oloc.argtypes = [P, P, dp]
        L.om_measure.argtypes = [P, C.c_int, ip, C.c_double, C.c_int, dp, dp, dp,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.om_measure_energy.argtypes = [P, C.c_int, ip, C.c_double, dp]
        L.om_bond_sdots.restype = C.c_double; L.om_bond_sdots.argtypes = [P, P, dp, dp]
        L.om_local_correlation.argtypes = [P, P, dp]
        L.om_measure_correlation.argtypes = [P, C.c_int, ip, C.c_double, dp, dp, dp]
        L.om_make_sample.argtypes = [P, P, P, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                     C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.om_sr_step.argtypes = [P, C.c_void_p, C.c_double, C.c_double, dp, dp, dp, dp, dp,
                                 dp]
        L.om_recompute.argtypes = [P, P, dp]
        L.om_bond.restype = C.c_int; L.om_bond.argtypes = [C.c_int] * 3
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Rng:
    def __init__(self, kind: str = "mt", seed: int = 1, stream: int = 0, counter: int = 0):
        L = lib()
        self.h = L.om_rng_mt(seed) if kind == "mt" else L.om_rng_philox(seed, stream, counter)

    def u01(self, n: int | None = None):
        if n is None:
            return lib().om_rng_u01(self.h)
        out = np.empty(n)
        lib().om_rng_fill(self.h, n, out)
        return out

    @property
    def counter(self) -> int:
        return int(lib().om_rng_counter(self.h))

    def __del__(self):
        try:
            lib().om_rng_free(self.h)
        except Exception:
            pass


def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().om_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def pfaffian(A: np.ndarray) -> float:
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().om_pfaffian(A.shape[0], np.ascontiguousarray(A.T.reshape(-1)))


class _SampleCfg(C.Structure):
    _fields_ = [("nsample", C.c_int), ("nwarmup", C.c_int), ("nupdate", C.c_int),
                ("fixed_k", C.c_int), ("ref_draws", C.c_int), ("p_restart", C.c_double),
                ("p_global", C.c_double), ("shift", C.c_double),
                ("cond_guard", C.c_double)]


class _SrCfg(C.Structure):
    _fields_ = [("s_factor0", C.c_double), ("delta_tau", C.c_double),
                ("max_change", C.c_double), ("slater_abs_av", C.c_double)]


class Oracle:
    """One model (tables + parameters) of the CPU restatement."""

    def __init__(self, spec, params=None):
        L = lib()
        self.spec = spec
        self.N = spec.N
        self.C = spec.C
        self.P = spec.P
        chi = np.ascontiguousarray(np.stack([spec.chi.real, spec.chi.imag], -1).reshape(-1))
        self.h = L.om_create(spec.L, spec.P, spec.alphac, spec.alphar, spec.n_slater,
                             spec.J1, spec.J2, spec.spin_parity,
                             np.ascontiguousarray(spec.w_map, np.int32).reshape(-1), chi,
                             np.ascontiguousarray(spec.orb, np.int32).reshape(-1))
        self.nv = L.om_nv(self.h)
        self.nv_rbm = L.om_nv_rbm(self.h)
        self.mc = spec.alphac * spec.P * 2
        self.mr = spec.alphar * spec.P * 2
        self.st = L.om_state_create(self.h)
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            lib().om_state_free(self.st)
            lib().om_free(self.h)
        except Exception:
            pass

    def set_params(self, p) -> None:
        def cplx(a):
            a = np.asarray(a, np.complex128).reshape(-1)
            return np.ascontiguousarray(np.stack([a.real, a.imag], -1).reshape(-1))
        rc = lib().om_set_params(self.h, cplx(p.Wc) if p.Wc.size else np.zeros(1),
                                 cplx(p.bc) if p.bc.size else np.zeros(1),
                                 np.ascontiguousarray(p.Wr, np.float64).reshape(-1)
                                 if p.Wr.size else np.zeros(1),
                                 np.ascontiguousarray(p.br, np.float64) if p.br.size else np.zeros(1),
                                 np.ascontiguousarray(p.f, np.float64))
        if rc:
            raise RuntimeError(f"om_set_params: {rc}")

    def set_gamma(self, gam) -> None:
        rc = lib().om_set_gamma(self.h, np.ascontiguousarray(gam, np.float64))
        if rc:
            raise RuntimeError(f"om_set_gamma: {rc}")

    def gamma(self) -> np.ndarray:
        out = np.zeros(self.nv)
        lib().om_get_gamma(self.h, out)
        return out

    def slater_elm(self) -> np.ndarray:
        n = self.C * 4 * self.N * self.N
        p = lib().om_slater_elm(self.h)
        return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(self.C, 2 * self.N, 2 * self.N).transpose(0, 2, 1)

    def set_projection(self, proj: int) -> None:
        """0: spin weights of normal_code; 1: of transverse_code
        (transverse_code/spin_projection.f90:30-38)."""
        rc = lib().om_set_projection(self.h, int(proj))
        if rc:
            raise RuntimeError(f"om_set_projection: {rc}")

    def sp_weights(self) -> np.ndarray:
        out = np.zeros(12)
        lib().om_sp_weights(self.h, out)
        return out.reshape(2, 6)

    def theta(self, x):
        thc = np.zeros(2 * max(self.mc, 1)); thr = np.zeros(max(self.mr, 1))
        lib().om_calc_theta(self.h, np.ascontiguousarray(x, np.int32), thc, thr)
        return thc[:2 * self.mc].view(np.complex128), thr[:self.mr]

    def eval(self, x, rng: Rng | None = None) -> dict:
        rc = lib().om_eval(self.h, np.ascontiguousarray(x, np.int32), self.st,
                           None if rng is None else rng.h)
        if rc:
            raise RuntimeError(f"om_eval: {rc}")
        return self.state()

    def state(self) -> dict:
        N, Cc = self.N, self.C
        x = np.zeros(N, np.int32); thc = np.zeros(2 * max(self.mc, 1)); thr = np.zeros(max(self.mr, 1))
        ei = np.zeros(N, np.int32); ec = np.zeros(N, np.int32); sg = np.zeros(self.P, np.int32)
        pf = np.zeros(Cc); invm = np.zeros(Cc * N * N); pwo = np.zeros(2 * Cc)
        psi = np.zeros(2 * (Cc + 1))
        lib().om_state_get(self.h, self.st, *[_ptr(a) for a in (x, thc, thr, ei, ec, sg, pf, invm, pwo, psi)])
        return dict(x=x, thetac=thc[:2 * self.mc].view(np.complex128), thetar=thr[:self.mr],
                    ele_idx=ei, ele_cfg=ec, sgn=sg, pf=pf,
                    invm=invm.reshape(Cc, N, N).transpose(0, 2, 1),  # [q, i, j] = InvM(i,j,q)
                    psi_wo=pwo.view(np.complex128), psi=psi.view(np.complex128))

    def recompute(self, state=None):
        """Full recompute of Pf and X^-1 in the state's electron order;
        returns the drift {max |dPf| / <|Pf|>, max |dInvM| / max |InvM|} of the
        fast-updated values it replaces."""
        drift = np.zeros(2)
        rc = lib().om_recompute(self.h, self.st if state is None else state, drift)
        if rc:
            raise RuntimeError(f"om_recompute: {rc}")
        return drift

    def ratio_exchange(self, i1, i2):
        ratio = np.zeros(self.C); pf = np.zeros(self.C); psi = np.zeros(2 * (self.C + 1))
        rc = lib().om_ratio_exchange(self.h, self.st, i1, i2, ratio, pf, psi)
        if rc:
            raise RuntimeError(f"om_ratio_exchange: {rc}")
        return ratio, pf, psi.view(np.complex128)

    def accept_exchange(self, i1, i2):
        rc = lib().om_accept_exchange(self.h, self.st, i1, i2, None)
        if rc:
            raise RuntimeError(f"om_accept_exchange: {rc}")

    def eloc(self) -> complex:
        out = np.zeros(2)
        lib().om_eloc(self.h, self.st, out, None)
        return complex(out[0], out[1])

    def oloc(self) -> np.ndarray:
        out = np.zeros(2 * self.nv)
        lib().om_oloc(self.h, self.st, out)
        return out.view(np.complex128)

    def measure(self, xs, shift=0.0, want_s=True, want_ostore=False) -> dict:
        xs = np.ascontiguousarray(xs, np.int32)
        ns = xs.shape[0]
        sums = np.zeros(4); g = np.zeros(self.nv); o = np.zeros(self.nv)
        S = np.zeros(self.nv * self.nv) if want_s else None
        el = np.zeros(2 * ns); ps = np.zeros(2 * ns)
        ost = np.zeros(2 * ns * self.nv) if want_ostore else None
        rc = lib().om_measure(self.h, ns, xs.reshape(-1), shift, int(want_s), sums, g, o,
                              _ptr(S), _ptr(el), _ptr(ps), _ptr(ost))
        if rc:
            raise RuntimeError(f"om_measure: {rc}")
        out = dict(Etot=sums[0], Etot2=sums[1], norm=sums[2], wf_av=sums[3], gvec=g, ovec=o,
                   eloc=el.view(np.complex128), psi=ps.view(np.complex128))
        if S is not None:
            out["smat"] = S.reshape(self.nv, self.nv).T.copy()
        if ost is not None:
            out["ostore"] = ost.reshape(2 * ns, self.nv)
        return out

    def measure_energy(self, xs, shift=0.0):
        xs = np.ascontiguousarray(xs, np.int32)
        sums = np.zeros(4)
        rc = lib().om_measure_energy(self.h, xs.shape[0], xs.reshape(-1), shift, sums)
        if rc:
            raise RuntimeError(f"om_measure_energy: {rc}")
        return dict(Etot=sums[0], Etot2=sums[1], norm=sums[2], wf_av=sums[3])

    def bond_sdots(self):
        """(E_loc, SdotS[4, N]) of the evaluated configuration."""
        e = np.zeros(2); sd = np.zeros(8 * self.N)
        lib().om_bond_sdots(self.h, self.st, e, sd)
        return complex(e[0], e[1]), sd.view(np.complex128).reshape(4, self.N)

    def local_correlation(self) -> np.ndarray:
        """Chi_loc_x[Nchi, 2] (SzSz, exchange) of the evaluated configuration."""
        nchi = self.N * (self.N + 1) // 2
        out = np.zeros(4 * nchi)
        lib().om_local_correlation(self.h, self.st, out)
        return out.view(np.complex128).reshape(nchi, 2)

    def measure_correlation(self, xs, shift=0.0) -> dict:
        """Un-normalised sums of the correlation run over the samples."""
        xs = np.ascontiguousarray(xs, np.int32)
        nchi = self.N * (self.N + 1) // 2
        sums = np.zeros(6); sp = np.zeros(2 * nchi); vb = np.zeros(4 * nchi)
        rc = lib().om_measure_correlation(self.h, xs.shape[0], xs.reshape(-1), shift, sums, sp, vb)
        if rc:
            raise RuntimeError(f"om_measure_correlation: {rc}")
        return dict(Etot=sums[0], Etot2=sums[1], norm=sums[2], wf_av=sums[3],
                    wf_sgn=complex(sums[4], sums[5]), chi_sp=sp.reshape(nchi, 2),
                    chi_vbs=vb.reshape(4, nchi))

    def make_sample(self, rng: Rng, first: bool, nsample: int, nwarmup: int,
                    nupdate: int | None = None, fixed_k: int = 0, ref_draws: bool = True,
                    p_restart: float = 0.05, p_global: float = 3e-4, shift: float = 0.0,
                    nfastup: int | None = None, state=None, cond_guard: float = 0.0):
        cfg = _SampleCfg(nsample, nwarmup, self.N if nupdate is None else nupdate, fixed_k,
                         int(ref_draws), p_restart, p_global, shift, cond_guard)
        nf = C.c_int(min(100, 2 * self.N) if nfastup is None else nfastup)
        xs = np.zeros((nsample, self.N), np.int32)
        psi = np.zeros(2 * nsample); pf = np.zeros(nsample * self.C)
        nacc = C.c_long(0); nrec = C.c_long(0)
        st = self.st if state is None else state
        rc = lib().om_make_sample(self.h, st, rng.h, int(first), C.byref(cfg), C.byref(nf),
                                  _ptr(xs), _ptr(psi), _ptr(pf), C.byref(nacc), C.byref(nrec))
        if rc:
            raise RuntimeError(f"om_make_sample: {rc}")
        return dict(x=xs, psi=psi.view(np.complex128), pf=pf.reshape(nsample, self.C),
                    naccept=nacc.value, nrecompute=nrec.value, nfastup=nf.value)

    def new_state(self):
        return lib().om_state_create(self.h)

    def sr_step(self, etot, norm, gvec, ovec, smat_raw, gamma, s_factor0=1e-3,
                delta_tau=0.05, max_change=0.04, slater_abs_av=0.25):
        cfg = _SrCfg(s_factor0, delta_tau, max_change, slater_abs_av)
        gam = np.ascontiguousarray(gamma, np.float64).copy()
        rep = np.zeros(8); x = np.zeros(self.nv)
        S = np.ascontiguousarray(np.asarray(smat_raw).T.reshape(-1))
        rc = lib().om_sr_step(self.h, C.byref(cfg), etot, norm,
                              np.ascontiguousarray(gvec, np.float64),
                              np.ascontiguousarray(ovec, np.float64), S, gam, rep, x)
        return rc, gam, rep, x
