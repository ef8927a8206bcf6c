"""ctypes front end of oracle/_ref/libref.so: the reference's own Fortran
routines (compiled from the reference sources by oracle/Makefile) behind the
bind(C) harness oracle/ref_harness.f90.  Test infrastructure only (tests/,
the golden-vector script and bench.py's cpu_baseline calibration); absent on
machines without the reference sources, where tests fall back to the
committed fixtures under tests/golden/."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libref.so")
# the same routines with transverse_code/spin_projection.f90
LIB_TC = os.path.join(HERE, "_ref", "libref_tc.so")
# the Calc_Correlation tree, with the correlation entry points
LIB_CC = os.path.join(HERE, "_ref", "libref_cc.so")


def _path(projection: int, tree: str | None) -> str:
    if tree == "cc":
        return LIB_CC
    return LIB_TC if projection else LIB

dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def available(projection: int = 0, tree: str | None = None) -> bool:
    return os.path.exists(_path(projection, tree))


class Ref:
    """One process-wide instance per library: the reference keeps its state
    in modules.  projection 1 loads the transverse_code build, tree "cc"
    the Calc_Correlation build."""

    _libs = {}

    def __init__(self, spec, params, projection: int = 0, tree: str | None = None):
        path = _path(projection, tree)
        if path not in Ref._libs:
            L = C.CDLL(path)
            L.ref_init.argtypes = [C.c_int] * 5 + [C.c_double, C.c_double, C.c_int, ip, dp, ip]
            L.ref_set_params.argtypes = [dp] * 5
            L.ref_get_slater_elm.argtypes = [dp]
            L.ref_get_sp.argtypes = [dp]
            L.ref_calc_theta.argtypes = [ip, dp, dp]
            L.ref_update_theta.argtypes = [ip, C.c_int, C.c_int, dp, dp]
            L.ref_x_to_ele.argtypes = [ip, ip, ip]
            L.ref_sign_pp.argtypes = [ip, ip]
            L.ref_amplitude.argtypes = [dp, dp, dp, ip, dp, dp]
            L.ref_pfm_two.argtypes = [C.c_int, C.c_int, ip, dp, dp, dp]
            L.ref_invm_two.argtypes = [C.c_int, C.c_int, ip, dp]
            L.ref_eloc.argtypes = [ip, dp, dp, dp, ip, ip, ip, dp, dp, dp]
            L.ref_deriv_pf.argtypes = [ip, dp, dp, dp]
            L.ref_bond_map.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int)]
            L.ref_sgrnd.argtypes = [C.c_int]
            L.ref_grnd_n.argtypes = [C.c_int, dp]
            if path == LIB_CC:
                L.ref_eloc_sdots.argtypes = [ip, dp, dp, dp, ip, ip, ip, dp, dp, dp, dp]
                L.ref_local_correlation.argtypes = [ip, dp, dp, dp, ip, ip, ip, dp, dp, dp]
            Ref._libs[path] = L
        self.L = Ref._libs[path]
        self.spec = spec
        s = spec
        chi = np.ascontiguousarray(np.stack([s.chi.real, s.chi.imag], -1).reshape(-1))
        self.L.ref_init(s.L, s.P, s.alphac, s.alphar, s.n_slater, s.J1, s.J2, s.spin_parity,
                        np.ascontiguousarray(s.w_map, np.int32).reshape(-1), chi,
                        np.ascontiguousarray(s.orb, np.int32).reshape(-1))
        self.mc = s.alphac * s.P * 2
        self.mr = s.alphar * s.P * 2
        self.set_params(params)

    def set_params(self, p):
        def cplx(a):
            a = np.asarray(a, np.complex128).reshape(-1)
            out = np.ascontiguousarray(np.stack([a.real, a.imag], -1).reshape(-1))
            return out if out.size else np.zeros(2)
        f = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) if np.size(a) else np.zeros(1)
        self.L.ref_set_params(cplx(p.Wc), cplx(p.bc), f(p.Wr), f(p.br), f(p.f))

    def slater_elm(self):
        N, Cc = self.spec.N, self.spec.C
        out = np.zeros(4 * N * N * Cc)
        self.L.ref_get_slater_elm(out)
        return out.reshape(Cc, 2 * N, 2 * N).transpose(0, 2, 1)

    def sp(self):
        out = np.zeros(12)
        self.L.ref_get_sp(out)
        return out.reshape(2, 6)

    def theta(self, x):
        thc = np.zeros(2 * max(self.mc, 1)); thr = np.zeros(max(self.mr, 1))
        self.L.ref_calc_theta(np.ascontiguousarray(x, np.int32), thc, thr)
        return thc[:2 * self.mc].view(np.complex128), thr[:self.mr]

    def update_theta(self, x, i1, i2, thc, thr):
        a = np.zeros(2 * max(self.mc, 1)); a[:2 * self.mc] = np.asarray(thc).view(np.float64)
        b = np.zeros(max(self.mr, 1)); b[:self.mr] = thr
        self.L.ref_update_theta(np.ascontiguousarray(x, np.int32), i1, i2, a, b)
        return a[:2 * self.mc].view(np.complex128), b[:self.mr]

    def x_to_ele(self, x):
        N = self.spec.N
        idx = np.zeros(N, np.int32); cfg = np.zeros(N, np.int32)
        self.L.ref_x_to_ele(np.ascontiguousarray(x, np.int32), idx, cfg)
        return idx, cfg

    def sign_pp(self, idx):
        out = np.zeros(self.spec.P, np.int32)
        self.L.ref_sign_pp(np.ascontiguousarray(idx, np.int32), out)
        return out

    def amplitude(self, thc, thr, pf, sgn):
        Cc = self.spec.C
        a = np.zeros(2 * max(self.mc, 1)); a[:2 * self.mc] = np.asarray(thc).view(np.float64)
        b = np.zeros(max(self.mr, 1)); b[:self.mr] = thr
        pwo = np.zeros(2 * Cc); psi = np.zeros(2 * (Cc + 1))
        self.L.ref_amplitude(a, b, np.ascontiguousarray(pf, np.float64),
                             np.ascontiguousarray(sgn, np.int32), pwo, psi)
        return pwo.view(np.complex128), psi.view(np.complex128)

    @staticmethod
    def _invm(invm):  # [q, i, j] = InvM(i,j,q) -> Fortran order buffer
        return np.ascontiguousarray(np.asarray(invm).transpose(0, 2, 1).reshape(-1))

    def pfm_two(self, msa, msb, tmp_idx, invm, pf_old):
        out = np.zeros(self.spec.C)
        self.L.ref_pfm_two(msa, msb, np.ascontiguousarray(tmp_idx, np.int32), self._invm(invm),
                           np.ascontiguousarray(pf_old, np.float64), out)
        return out

    def invm_two(self, msa, msb, idx, invm):
        N, Cc = self.spec.N, self.spec.C
        buf = self._invm(invm).copy()
        self.L.ref_invm_two(msa, msb, np.ascontiguousarray(idx, np.int32), buf)
        return buf.reshape(Cc, N, N).transpose(0, 2, 1)

    def eloc(self, st):
        out = np.zeros(2)
        thc = np.zeros(2 * max(self.mc, 1)); thc[:2 * self.mc] = st["thetac"].view(np.float64)
        thr = np.zeros(max(self.mr, 1)); thr[:self.mr] = st["thetar"]
        psi0 = np.array([st["psi"][0].real, st["psi"][0].imag])
        self.L.ref_eloc(np.ascontiguousarray(st["x"], np.int32), psi0, thc, thr,
                        np.ascontiguousarray(st["ele_idx"], np.int32),
                        np.ascontiguousarray(st["ele_cfg"], np.int32),
                        np.ascontiguousarray(st["sgn"], np.int32),
                        np.ascontiguousarray(st["pf"], np.float64), self._invm(st["invm"]), out)
        return complex(out[0], out[1])

    def deriv_pf(self, st):
        Cc, nsl = self.spec.C, self.spec.n_slater
        out = np.zeros(nsl * Cc)
        self.L.ref_deriv_pf(np.ascontiguousarray(st["ele_idx"], np.int32),
                            np.ascontiguousarray(st["pf"], np.float64), self._invm(st["invm"]), out)
        return out.reshape(Cc, nsl)

    def bond(self, t, i):
        o = C.c_int(0)
        self.L.ref_bond_map(t, i, C.byref(o))
        return o.value

    def grnd(self, seed, n):
        self.L.ref_sgrnd(seed)
        out = np.zeros(n)
        self.L.ref_grnd_n(n, out)
        return out

    def _state_args(self, st):
        thc = np.zeros(2 * max(self.mc, 1)); thc[:2 * self.mc] = st["thetac"].view(np.float64)
        thr = np.zeros(max(self.mr, 1)); thr[:self.mr] = st["thetar"]
        psi0 = np.array([st["psi"][0].real, st["psi"][0].imag])
        return (np.ascontiguousarray(st["x"], np.int32), psi0, thc, thr,
                np.ascontiguousarray(st["ele_idx"], np.int32), np.ascontiguousarray(st["ele_cfg"], np.int32),
                np.ascontiguousarray(st["sgn"], np.int32), np.ascontiguousarray(st["pf"], np.float64),
                self._invm(st["invm"]))

    def eloc_sdots(self, st):
        """Calc_Correlation tree: (E_loc, SdotS[4, N])."""
        e = np.zeros(2); sd = np.zeros(8 * self.spec.N)
        self.L.ref_eloc_sdots(*self._state_args(st), e, sd)
        return complex(e[0], e[1]), sd.view(np.complex128).reshape(4, self.spec.N)

    def local_correlation(self, st):
        """Calc_Correlation tree: Chi_loc_x as [Nchi, 2]."""
        N = self.spec.N
        out = np.zeros(2 * N * (N + 1))
        self.L.ref_local_correlation(*self._state_args(st), out)
        return out.view(np.complex128).reshape(N * (N + 1) // 2, 2)
