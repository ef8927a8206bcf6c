"""Python mirror of the reference driver's calls into the VMC hot path,
over the C ABI of include/rbmpp.h (librbmpp.so, gfx950).

Method names follow the reference routines they replace:
  update_slater_elements / set_params  -> slater.f90:7-64 (+ BCAST main.f90:251-259)
  make_sample                          -> make_sample.f90:4-240
  calc_phys_quantity_opt               -> calc_phys_quantity.f90:20-246 + main.f90:347-367
  sr_step                              -> SR_pdposv.f90:1-414
  calc_phys_quantity_after_opt         -> calc_phys_quantity.f90:250-302
Errors from the library raise RbmppError; there is no CPU fallback: if the
HIP library is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import lattice as lt

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librbmpp.so")

S_FP64, S_FP32, S_BF16X3 = 0, 1, 2
PROJ_NORMAL, PROJ_TRANSVERSE, PROJ_NORMAL_EXACT = 0, 1, 2
SR_NO_GRAPH = 1
SR_ASSEMBLED_REFINE = 2


TUNE_S_BUILD, TUNE_MEASURE = 1, 2


class TuneReport(C.Structure):
    _fields_ = [("s_tile", C.c_int32), ("s_nsplit", C.c_int32), ("s_default_tile", C.c_int32),
                ("s_default_nsplit", C.c_int32), ("s_build_s", C.c_double), ("measure_grid", C.c_int32),
                ("measure_configurations", C.c_int32), ("measure_s", C.c_double)]


class RbmppError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rbmpp error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("L", C.c_int32), ("n_sym", C.c_int32), ("alphac", C.c_int32), ("alphar", C.c_int32),
                ("n_slater", C.c_int32), ("spin_parity", C.c_int32), ("J1", C.c_double),
                ("J2", C.c_double), ("n_walkers", C.c_int32), ("n_samples", C.c_int32),
                ("n_warmup", C.c_int32), ("n_warmup_first", C.c_int32), ("n_update", C.c_int32),
                ("recompute_k", C.c_int32), ("s_precision", C.c_int32), ("fused_measure", C.c_int32),
                ("seed", C.c_uint64), ("p_restart", C.c_double), ("p_global", C.c_double),
                ("cond_guard", C.c_double)]


class SrConfig(C.Structure):
    _fields_ = [("s_factor0", C.c_double), ("delta_tau", C.c_double), ("max_change", C.c_double),
                ("slater_abs_av", C.c_double), ("refine_iters", C.c_int32), ("flags", C.c_int32)]


class SrReport(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("delw_max", C.c_double), ("scaling", C.c_double),
                ("sdiag_av", C.c_double), ("s_factor", C.c_double), ("residual", C.c_double),
                ("mean_abs_f", C.c_double), ("k_nonzero", C.c_int32), ("cg_iters", C.c_int32),
                ("precond_boost", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.rbmpp_create.argtypes = [C.POINTER(Config), ip, dp, ip, C.POINTER(vp)]
        L.rbmpp_destroy.argtypes = [vp]
        L.rbmpp_last_error.restype = C.c_char_p; L.rbmpp_last_error.argtypes = [vp]
        L.rbmpp_nv.argtypes = [vp]
        L.rbmpp_set_params.argtypes = [vp, dp, dp, dp, dp, dp]
        L.rbmpp_set_gamma.argtypes = [vp, dp]
        L.rbmpp_get_gamma.argtypes = [vp, dp]
        L.rbmpp_make_sample.argtypes = [vp, C.c_int32, C.c_double, C.POINTER(C.c_int64)]
        L.rbmpp_measure.argtypes = [vp, C.c_double] + [dp] * 6
        L.rbmpp_sr_step.argtypes = [vp, C.POINTER(SrConfig), dp, C.POINTER(SrReport)]
        L.rbmpp_measure_energy.argtypes = [vp, C.c_double] + [dp] * 4
        L.rbmpp_measure_correlation.argtypes = [vp, C.c_double] + [dp] * 7
        L.rbmpp_set_samples.argtypes = [vp, C.c_int32, ip]
        L.rbmpp_get_samples.argtypes = [vp, ip, dp]
        L.rbmpp_get_measured.argtypes = [vp, dp, dp, dp]
        L.rbmpp_set_measured.argtypes = [vp, C.c_int64, dp, dp, dp, dp, C.c_uint64]
        L.rbmpp_n_samples.restype = C.c_int64; L.rbmpp_n_samples.argtypes = [vp]
        L.rbmpp_get_walkers.argtypes = [vp, ip, dp, dp, dp, C.POINTER(C.c_int64)]
        L.rbmpp_get_walker_electrons.argtypes = [vp, ip]
        L.rbmpp_eval.argtypes = [vp, C.c_int32, ip, dp, dp, dp]
        L.rbmpp_ratio_exchange.argtypes = [vp, C.c_int32, ip, ip, ip, dp, dp]
        L.rbmpp_local.argtypes = [vp, C.c_int32, ip, dp, dp, dp]
        L.rbmpp_local_noalloc.argtypes = [vp, C.c_int32, ip, dp, dp]
        L.rbmpp_get_smat.argtypes = [vp, dp]
        L.rbmpp_timings.argtypes = [vp, dp]
        L.rbmpp_set_s_precision.argtypes = [vp, C.c_int32]
        L.rbmpp_set_projection.argtypes = [vp, C.c_int32]
        L.rbmpp_tune.argtypes = [vp, C.c_int32, C.POINTER(TuneReport)]
        L.rbmpp_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.rbmpp_batch_make_sample.argtypes = [vp, ip, dp, C.POINTER(C.c_int64)]
        L.rbmpp_batch_sample_seconds.restype = C.c_double
        L.rbmpp_batch_sample_seconds.argtypes = [vp]
        L.rbmpp_batch_last_error.restype = C.c_char_p
        L.rbmpp_batch_last_error.argtypes = [vp]
        L.rbmpp_batch_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


class Engine:
    """One model on one GPU: tables, parameters, walkers, samples, S."""

    def __init__(self, spec: lt.ModelSpec, n_walkers: int, n_samples: int, seed: int = 1,
                 n_warmup: int = -1, n_warmup_first: int = -1, n_update: int = 0,
                 recompute_k: int = 0, s_precision: int = S_FP64, p_restart: float = 0.05,
                 p_global: float = 3e-4, params: lt.Params | None = None,
                 fused_measure: bool = False, cond_guard: float = 1e4):
        self.spec = spec
        N, P = spec.N, spec.P
        self.cfg = Config(spec.L, P, spec.alphac, spec.alphar, spec.n_slater, spec.spin_parity,
                          spec.J1, spec.J2, n_walkers, n_samples, n_warmup, n_warmup_first, n_update,
                          recompute_k, s_precision, int(fused_measure), seed, p_restart, p_global,
                          cond_guard)
        # reference layouts: W_map(i,1,g) 1-based; Orbital_Idx(i,j) column-major 1-based
        wmap = np.ascontiguousarray(spec.w_map.astype(np.int32) + 1)
        chi = np.ascontiguousarray(np.stack([spec.chi.real, spec.chi.imag], -1).reshape(-1))
        orb = np.ascontiguousarray(spec.orb.T.astype(np.int32) + 1)
        h = C.c_void_p()
        rc = lib().rbmpp_create(C.byref(self.cfg), _i(wmap), _d(chi), _i(orb), C.byref(h))
        self.h = h
        self._check(rc)
        self.nv = spec.nv
        self.C = spec.C
        self.N = N
        self.W = n_walkers
        self.nspw = n_samples
        if params is not None:
            self.set_params(params)

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = lib().rbmpp_last_error(self.h).decode() if self.h else ""
            raise RbmppError(rc, msg)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().rbmpp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def set_params(self, p: lt.Params) -> None:
        cw = lambda a: np.ascontiguousarray(np.stack([np.real(a), np.imag(a)], -1).reshape(-1)
                                            if np.size(a) else np.zeros(2))
        fw = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)
                                            if np.size(a) else np.zeros(1))
        self._check(lib().rbmpp_set_params(self.h, _d(cw(p.Wc)), _d(cw(p.bc)), _d(fw(p.Wr)),
                                           _d(fw(p.br)), _d(fw(p.f))))

    update_slater_elements = set_params

    def set_gamma(self, gam) -> None:
        g = np.ascontiguousarray(gam, np.float64)
        self._check(lib().rbmpp_set_gamma(self.h, _d(g)))

    def gamma(self) -> np.ndarray:
        out = np.zeros(self.nv)
        self._check(lib().rbmpp_get_gamma(self.h, _d(out)))
        return out

    # ---- hot path
    def make_sample(self, first_iteration: bool, shift: float = 0.0) -> int:
        n = C.c_int64(0)
        self._check(lib().rbmpp_make_sample(self.h, int(first_iteration), shift, C.byref(n)))
        return n.value

    def calc_phys_quantity_opt(self, shift: float = 0.0) -> dict:
        e, e2, nr, wf = (C.c_double() for _ in range(4))
        g = np.zeros(self.nv); o = np.zeros(self.nv)
        self._check(lib().rbmpp_measure(self.h, shift, C.byref(e), C.byref(e2), C.byref(nr),
                                        C.byref(wf), _d(g), _d(o)))
        return dict(Etot=e.value, Etot2=e2.value, norm=nr.value, wf_av=wf.value, gvec=g, ovec=o)

    def calc_phys_quantity_after_opt(self, shift: float = 0.0) -> dict:
        e, e2, nr, wf = (C.c_double() for _ in range(4))
        self._check(lib().rbmpp_measure_energy(self.h, shift, C.byref(e), C.byref(e2), C.byref(nr),
                                               C.byref(wf)))
        return dict(Etot=e.value, Etot2=e2.value, norm=nr.value, wf_av=wf.value)

    def measure_correlation(self, shift: float = 0.0) -> dict:
        """The correlation run's measurement on the stored samples
        (Calc_Correlation/calc_phys_quantity.f90:264-550 with the
        normalisation of Calc_Correlation/main.f90:543-581): chi_sp[i0, k] =
        Chi_sp(k+1, i0+1), chi_vbs[t, i0] = Chi_vbs(i0+1, t+1), pairs
        i1 <= i2 numbered i0 = i2 (i2 + 1) / 2 + i1 (0-based)."""
        N = self.spec.N
        nchi = N * (N + 1) // 2
        e, e2, nr, wf = (C.c_double() for _ in range(4))
        ws = np.zeros(2); sp = np.zeros(2 * nchi); vb = np.zeros(4 * nchi)
        self._check(lib().rbmpp_measure_correlation(self.h, shift, C.byref(e), C.byref(e2), C.byref(nr),
                                                    C.byref(wf), _d(ws), _d(sp), _d(vb)))
        return dict(Etot=e.value, Etot2=e2.value, norm=nr.value, wf_av=wf.value,
                    wf_sgn=complex(ws[0], ws[1]), chi_sp=sp.reshape(nchi, 2), chi_vbs=vb.reshape(4, nchi))

    def sr_step(self, gamma=None, s_factor0: float = 1e-3, delta_tau: float = 0.05,
                max_change: float = 0.04, slater_abs_av: float = 0.25, refine_iters: int = 1,
                graph: bool = True, assembled_refine: bool = False):
        """graph=False runs the SR system's launches eagerly instead of
        replaying the captured graph (RBMPP_SR_NO_GRAPH); assembled_refine
        refines a mixed-precision S against the assembled matrix instead of
        the FP64 operator of the O rows (RBMPP_SR_ASSEMBLED_REFINE)."""
        g = np.ascontiguousarray(self.gamma() if gamma is None else gamma, np.float64).copy()
        flags = (0 if graph else SR_NO_GRAPH) | (SR_ASSEMBLED_REFINE if assembled_refine else 0)
        c = SrConfig(s_factor0, delta_tau, max_change, slater_abs_av, refine_iters, flags)
        rep = SrReport()
        self._check(lib().rbmpp_sr_step(self.h, C.byref(c), _d(g), C.byref(rep)))
        return g, {k: getattr(rep, k) for k, _ in SrReport._fields_}

    # ---- parity / inspection entry points
    def set_samples(self, xs) -> None:
        xs = np.ascontiguousarray(xs, np.int32)
        self._check(lib().rbmpp_set_samples(self.h, xs.shape[0], _i(xs)))

    def get_samples(self):
        n = lib().rbmpp_n_samples(self.h)
        x = np.zeros((n, self.N), np.int32); psi = np.zeros(2 * n)
        self._check(lib().rbmpp_get_samples(self.h, _i(x), _d(psi)))
        return x, psi.view(np.complex128)

    def get_measured(self):
        """Per-sample psi, reweighted E_loc and reweight factor of the last
        measurement."""
        n = lib().rbmpp_n_samples(self.h)
        psi = np.zeros(2 * n); el = np.zeros(2 * n); rw = np.zeros(n)
        self._check(lib().rbmpp_get_measured(self.h, _d(psi), _d(el), _d(rw)))
        return psi.view(np.complex128), el.view(np.complex128), rw

    def set_measured(self, n: int, psi=None, eloc=None, rw=None, orows=None, seed: int = 1) -> None:
        """Load a measurement (psi, reweighted eloc, rw per configuration and
        the O rows) in place of calc_phys_quantity_opt; with orows=None a
        seeded synthetic one is generated on the device."""
        if orows is None:
            self._check(lib().rbmpp_set_measured(self.h, n, None, None, None, None, seed))
        else:
            c2 = lambda a: np.ascontiguousarray(np.stack([np.real(a), np.imag(a)], -1).reshape(-1))
            o = np.ascontiguousarray(orows, np.float64)
            self._check(lib().rbmpp_set_measured(self.h, n, _d(c2(psi)), _d(c2(eloc)),
                                                 _d(np.ascontiguousarray(rw, np.float64)), _d(o), seed))

    def get_walkers(self, with_invm: bool = False):
        W, N, Cc = self.W, self.N, self.C
        x = np.zeros((W, N), np.int32); pf = np.zeros((W, Cc)); psi = np.zeros(2 * W)
        inv = np.zeros(W * Cc * N * N) if with_invm else None
        nacc = np.zeros(W, np.int64)
        self._check(lib().rbmpp_get_walkers(self.h, _i(x), _d(pf), _d(psi), _d(inv),
                                            nacc.ctypes.data_as(C.POINTER(C.c_int64))))
        out = dict(x=x, pf=pf, psi=psi.view(np.complex128), naccept=nacc)
        if with_invm:
            out["invm"] = inv.reshape(W, Cc, N, N).transpose(0, 1, 3, 2)  # [w,q,i,j] = InvM(i,j)
        return out

    def walker_electrons(self) -> np.ndarray:
        """Ele_Idx of every walker, 0-based sites: [w, i] = site of electron i."""
        e = np.zeros((self.W, self.N), np.int32)
        self._check(lib().rbmpp_get_walker_electrons(self.h, _i(e)))
        return e - 1

    def eval(self, xs, with_invm: bool = False):
        xs = np.ascontiguousarray(np.atleast_2d(xs), np.int32)
        n = xs.shape[0]
        psi = np.zeros(n * 2 * (self.C + 1)); pf = np.zeros(n * self.C)
        inv = np.zeros(n * self.C * self.N * self.N) if with_invm else None
        self._check(lib().rbmpp_eval(self.h, n, _i(xs), _d(psi), _d(pf), _d(inv)))
        out = dict(psi=psi.view(np.complex128).reshape(n, self.C + 1), pf=pf.reshape(n, self.C))
        if with_invm:
            out["invm"] = inv.reshape(n, self.C, self.N, self.N).transpose(0, 1, 3, 2)
        return out

    def ratio_exchange(self, xs, i1, i2):
        xs = np.ascontiguousarray(np.atleast_2d(xs), np.int32)
        n = xs.shape[0]
        a = np.ascontiguousarray(i1, np.int32); b = np.ascontiguousarray(i2, np.int32)
        pf = np.zeros(n * self.C); psi = np.zeros(2 * n)
        self._check(lib().rbmpp_ratio_exchange(self.h, n, _i(xs), _i(a), _i(b), _d(pf), _d(psi)))
        return pf.reshape(n, self.C), psi.view(np.complex128)

    def local(self, xs, want_o: bool = True):
        xs = np.ascontiguousarray(np.atleast_2d(xs), np.int32)
        n = xs.shape[0]
        el = np.zeros(2 * n); psi = np.zeros(2 * n)
        o = np.zeros(2 * n * self.nv) if want_o else None
        self._check(lib().rbmpp_local(self.h, n, _i(xs), _d(el), _d(psi), _d(o)))
        out = dict(eloc=el.view(np.complex128), psi=psi.view(np.complex128))
        if want_o:
            out["oloc"] = o.view(np.complex128).reshape(n, self.nv)
        return out

    def local_noalloc(self, xs):
        """rbmpp_local without O through the allocation-free entry point
        (device staging in the library's fixed array)."""
        xs = np.ascontiguousarray(np.atleast_2d(xs), np.int32)
        n = xs.shape[0]
        el = np.zeros(2 * n); psi = np.zeros(2 * n)
        self._check(lib().rbmpp_local_noalloc(self.h, n, _i(xs), _d(el), _d(psi)))
        return dict(eloc=el.view(np.complex128), psi=psi.view(np.complex128))

    def tune(self, what: int = TUNE_S_BUILD) -> dict:
        """rbmpp_tune: time the S build's tile / split-K choices and (or) the
        estimator grid on this engine's data, keep the fastest."""
        r = TuneReport()
        self._check(lib().rbmpp_tune(self.h, what, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneReport._fields_}

    def set_projection(self, projection: int) -> None:
        """Spin-projection weights: PROJ_NORMAL (normal_code, cos(pi/2) taken as
        0: block form), PROJ_NORMAL_EXACT (normal_code with the reference's
        cos(pi/2) = 6.1e-17: general skew form) or PROJ_TRANSVERSE
        (transverse_code/spin_projection.f90:30-38); rebuilds Slater_Elm."""
        self._check(lib().rbmpp_set_projection(self.h, int(projection)))

    def set_s_precision(self, prec: int) -> None:
        self._check(lib().rbmpp_set_s_precision(self.h, prec))

    def smat(self) -> np.ndarray:
        S = np.zeros(self.nv * self.nv)
        self._check(lib().rbmpp_get_smat(self.h, _d(S)))
        return S.reshape(self.nv, self.nv)

    def timings(self) -> dict:
        t = np.zeros(8)
        self._check(lib().rbmpp_timings(self.h, _d(t)))
        return dict(slater=t[0], sample=t[1], measure=t[2], s_build=t[3], solve=t[4], sr_total=t[5],
                    recompute_count=t[6])


class Batch:
    """Several engines (lattices, couplings, parameters) sampled by one
    persistent launch over all their walkers (rbmpp_batch_make_sample); each
    engine measures and takes its SR step on its own."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        rc = lib().rbmpp_batch_create(arr, len(self.engines), C.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().rbmpp_batch_last_error(h).decode() if h else ""
            self.close()
            raise RbmppError(rc, msg)

    def make_sample(self, first_iteration, shift) -> np.ndarray:
        n = len(self.engines)
        first = np.ascontiguousarray(np.broadcast_to(np.asarray(first_iteration, np.int32), (n,)))
        sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, np.float64), (n,)))
        nacc = np.zeros(n, np.int64)
        rc = lib().rbmpp_batch_make_sample(self.h, _i(first), _d(sh), nacc.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc != 0:
            raise RbmppError(rc, lib().rbmpp_batch_last_error(self.h).decode())
        return nacc

    def sample_seconds(self) -> float:
        return lib().rbmpp_batch_sample_seconds(self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().rbmpp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
