"""Lattice tables and the reference's input/output file formats.

The reference ships no inputs; its driver reads them from files
(main.f90:75-129, parameter.f90:42-141).  This module generates the tables
for an L x L periodic square lattice and reads/writes those files in the
reference's list-directed formats, so a run set up here can be handed to the
reference Fortran and vice versa.

Index conventions (0-based inside Python, 1-based in the files):
  site s = ix*L + iy           (calc_phys_quantity.f90:394-422, bond_map)
  W_map[g, i]  image of site i under symmetry operation g (sym_info.txt)
  orb[i, j]    orbital index of the pair amplitude f_ij (orbitalidx.def)
"""
from __future__ import annotations

import dataclasses
import os
from typing import Sequence

import numpy as np

# C4v operations about the site (0, 0), acting on lattice coordinates.
C4V_OPS = (
    lambda x, y: (x, y),
    lambda x, y: (-y, x),
    lambda x, y: (-x, -y),
    lambda x, y: (y, -x),
    lambda x, y: (-x, y),
    lambda x, y: (x, -y),
    lambda x, y: (y, x),
    lambda x, y: (-y, -x),
)


def site(ix: int, iy: int, L: int) -> int:
    return (ix % L) * L + (iy % L)


def c4v_maps(L: int, translations: bool = False) -> np.ndarray:
    """Site permutations of the C4v point group about (0,0), optionally
    combined with all L*L translations (momentum (0,0) projection).
    Returns W_map[P, N] with 0-based images."""
    N = L * L
    maps = []
    shifts = [(tx, ty) for tx in range(L) for ty in range(L)] if translations else [(0, 0)]
    for tx, ty in shifts:
        for op in C4V_OPS:
            m = np.empty(N, dtype=np.int32)
            for ix in range(L):
                for iy in range(L):
                    nx, ny = op(ix, iy)
                    m[site(ix, iy, L)] = site(nx + tx, ny + ty, L)
            maps.append(m)
    return np.array(maps, dtype=np.int32)


def sublattice_orbitals(L: int) -> tuple[np.ndarray, int]:
    """mVMC-style 2x2-sublattice pair orbitals: f_ij depends on the 2x2
    sublattice of i and the displacement r_j - r_i.  Returns (orb[N,N], 4N).
    The N_Slater/N = 4 displacement-zero orbitals are the f_ii that the SR
    step expects to carry zero derivative (SR_pdposv.f90:33,78-100)."""
    if L % 2:
        raise ValueError("2x2 sublattice orbitals need even L")
    N = L * L
    orb = np.empty((N, N), dtype=np.int32)
    for ix in range(L):
        for iy in range(L):
            i = site(ix, iy, L)
            sub = (ix % 2) * 2 + (iy % 2)
            for jx in range(L):
                for jy in range(L):
                    j = site(jx, jy, L)
                    disp = ((jx - ix) % L) * L + ((jy - iy) % L)
                    orb[i, j] = sub * N + disp
    return orb, 4 * N


def bonds(L: int) -> np.ndarray:
    """bond_map (calc_phys_quantity.f90:382-427): [4, N] partner sites for
    bond types (x+1,y), (x,y+1) [J1] and (x+1,y+1), (x+1,y-1) [J2]."""
    N = L * L
    out = np.empty((4, N), dtype=np.int32)
    for i in range(N):
        ix, iy = divmod(i, L)
        out[0, i] = site(ix + 1, iy, L)
        out[1, i] = site(ix, iy + 1, L)
        out[2, i] = site(ix + 1, iy + 1, L)
        out[3, i] = site(ix + 1, iy - 1, L)
    return out


def sublattice_b(L: int) -> np.ndarray:
    """define_sublattice (interface_for_PP.f90:110-136)."""
    N = L * L
    return np.array([((s // L) + (s % L)) & 1 for s in range(N)], dtype=np.int32)


@dataclasses.dataclass
class ModelSpec:
    """Everything the reference reads from RBM.input, sym_info.txt and
    orbitalidx.def (main.f90:75-141)."""
    L: int
    alphac: int
    alphar: int
    J1: float
    J2: float
    spin_parity: int
    w_map: np.ndarray          # [P, N] int32, 0-based
    chi: np.ndarray            # [P] complex128 (para_mp_trans)
    orb: np.ndarray            # [N, N] int32, 0-based
    n_slater: int
    nsample: int = 0
    nstep: int = 0
    nav: int = 0
    delta_tau: float = 0.05
    s_factor0: float = 1e-3
    s_factor_scale: float = 1.0
    s_cut: float = 0.0

    @property
    def N(self) -> int:
        return self.L * self.L

    @property
    def P(self) -> int:
        return int(self.w_map.shape[0])

    @property
    def C(self) -> int:
        return 2 * self.P

    @property
    def nv(self) -> int:
        N = self.N
        return self.alphac * 2 * (N + 1) + self.alphar * (N + 1) + self.n_slater

    @property
    def nv_rbm(self) -> int:
        return self.nv - self.n_slater


def j1j2_model(L: int, alphar: int, J2: float, J1: float = 1.0, alphac: int = 0,
               spin_parity: int = 1, translations: bool = False) -> ModelSpec:
    w_map = c4v_maps(L, translations)
    orb, nsl = sublattice_orbitals(L)
    chi = np.ones(w_map.shape[0], dtype=np.complex128)
    return ModelSpec(L=L, alphac=alphac, alphar=alphar, J1=J1, J2=J2,
                     spin_parity=spin_parity, w_map=w_map, chi=chi, orb=orb,
                     n_slater=nsl)


@dataclasses.dataclass
class Params:
    Wc: np.ndarray  # [alphac, N] complex128  (Wcirr(iw,f) -> Wc[f, iw])
    bc: np.ndarray  # [alphac] complex128
    Wr: np.ndarray  # [alphar, N] float64
    br: np.ndarray  # [alphar] float64
    f: np.ndarray   # [N_Slater] float64

    def gamma(self) -> np.ndarray:
        """Pack in the order of main.f90:263-294."""
        parts = []
        for b in self.bc:
            parts += [b.real, b.imag]
        for w in self.Wc.reshape(-1):
            parts += [w.real, w.imag]
        out = np.concatenate([np.array(parts, dtype=np.float64),
                              self.br.astype(np.float64),
                              self.Wr.reshape(-1).astype(np.float64),
                              self.f.astype(np.float64)])
        return out

    @staticmethod
    def from_gamma(spec: ModelSpec, gam: np.ndarray) -> "Params":
        """Unpack as SR_pdposv.f90:385-412."""
        N, ac, ar = spec.N, spec.alphac, spec.alphar
        k = 0
        bc = gam[k:k + 2 * ac].reshape(ac, 2); k += 2 * ac
        Wc = gam[k:k + 2 * ac * N].reshape(ac, N, 2); k += 2 * ac * N
        br = gam[k:k + ar].copy(); k += ar
        Wr = gam[k:k + ar * N].reshape(ar, N).copy(); k += ar * N
        f = gam[k:k + spec.n_slater].copy()
        return Params(Wc=Wc[..., 0] + 1j * Wc[..., 1], bc=bc[:, 0] + 1j * bc[:, 1],
                      Wr=Wr, br=br, f=f)


def renormalise_f(f: np.ndarray, slater_abs_av: float = 0.25) -> np.ndarray:
    """parameter.f90:121-140: scale f so that mean|f| = slater_abs_av, with the
    reference's arithmetic (sequential sum of |f|, then f * slater_abs_av / mean)
    so that the result is bit-identical to its reader's."""
    f = np.asarray(f, np.float64)
    r3 = 0.0
    for v in f.tolist():
        r3 += abs(v)
    r3 /= float(len(f))
    return f * slater_abs_av / r3


def random_params(spec: ModelSpec, seed_w: int = 1, seed_f: int = 2,
                  sigma_w: float = 0.01) -> Params:
    """W, b ~ N(0, sigma_w^2) (seed_w); f ~ N(0,1) (seed_f) renormalised to
    mean|f| = 0.25.  The reference instead draws W uniformly in
    [-0.05, 0.05) from a clock-seeded stream (parameter.f90:21-41)."""
    rw = np.random.default_rng(seed_w)
    N = spec.N
    Wc = (rw.normal(0, sigma_w, (spec.alphac, N))
          + 1j * rw.normal(0, sigma_w, (spec.alphac, N)))
    bc = rw.normal(0, sigma_w, spec.alphac) + 1j * rw.normal(0, sigma_w, spec.alphac)
    Wr = rw.normal(0, sigma_w, (spec.alphar, N))
    br = rw.normal(0, sigma_w, spec.alphar)
    rf = np.random.default_rng(seed_f)
    f = renormalise_f(rf.normal(0, 1, spec.n_slater))
    return Params(Wc=Wc, bc=bc, Wr=Wr, br=br, f=f)


# ---------------------------------------------------------------------------
# reference file formats
# ---------------------------------------------------------------------------
def _data_lines(path: str) -> list[str]:
    with open(path) as fh:
        return fh.read().splitlines()


def write_rbm_input(path: str, spec: ModelSpec) -> None:
    """RBM.input (main.f90:75-84)."""
    with open(path, "w") as fh:
        fh.write("L Nsample spin_parity\n")
        fh.write(f"{spec.L} {spec.nsample} {spec.spin_parity}\n")
        fh.write("alphac alphar Nstep Nav delta_tau\n")
        fh.write(f"{spec.alphac} {spec.alphar} {spec.nstep} {spec.nav} {spec.delta_tau!r}\n")
        fh.write("s_factor0 s_factor_scale s_cut\n")
        fh.write(f"{spec.s_factor0!r} {spec.s_factor_scale!r} {spec.s_cut!r}\n")
        fh.write("J1 J2\n")
        fh.write(f"{spec.J1!r} {spec.J2!r}\n")


def write_sym_info(path: str, w_map: np.ndarray, chi: Sequence[complex]) -> None:
    """sym_info.txt (main.f90:89-124): complex factors list-directed as (re,im)."""
    P, N = w_map.shape
    with open(path, "w") as fh:
        fh.write("# N_MP_Trans Ncopy\n")
        fh.write(f"{P} 1\n")
        fh.write("# mpidx para_mp_trans\n")
        for g in range(P):
            c = complex(chi[g])
            fh.write(f"{g + 1} ({c.real!r},{c.imag!r})\n")
        fh.write("# mpidx icopy i W_map(i)\n")
        for g in range(P):
            for i in range(N):
                fh.write(f"{g + 1} 1 {i + 1} {int(w_map[g, i]) + 1}\n")


def write_orbitalidx(path: str, orb: np.ndarray, n_slater: int) -> None:
    """orbitalidx.def (mVMC style; parameter.f90:81-101)."""
    N = orb.shape[0]
    with open(path, "w") as fh:
        fh.write("=============================================\n")
        fh.write(f"NOrbitalIdx {n_slater}\n")
        fh.write("ComplexType 0\n")
        fh.write("=============================================\n")
        fh.write("=============================================\n")
        for i in range(N):
            for j in range(N):
                fh.write(f"{i} {j} {int(orb[i, j])}\n")


def write_zqp_orbital(path: str, f: np.ndarray) -> None:
    """zqp_orbital_initial.dat / zqp_orbital_opt.dat (parameter.f90:114-141,
    266-275)."""
    with open(path, "w") as fh:
        fh.write("\n")
        fh.write(f" Slater {len(f):10d}\n")
        fh.write("\n\n\n")
        for k, v in enumerate(f):
            fh.write(f"{k:5d}{v:25.15e}{0.0:25.15e}\n")


def write_w(path: str, p: Params) -> None:
    """W.input / optimized_W.txt (parameter.f90:248-264)."""
    with open(path, "w") as fh:
        fh.write("# index for independent neurons (f), index for interaction (iw), "
                 "irreducible interaction Wirr(iw,f)\n")
        for f in range(len(p.bc)):
            fh.write(f"# bias{f + 1:6d}{p.bc[f].real:22.15f}{p.bc[f].imag:22.15f}\n")
            for iw in range(p.Wc.shape[1]):
                w = p.Wc[f, iw]
                fh.write(f"{f + 1:6d}{iw + 1:6d}{w.real:22.15f}{w.imag:22.15f}\n")
            fh.write("\n")
        for f in range(len(p.br)):
            fh.write(f"# bias{f + 1:6d}{p.br[f]:22.15f}{0.0:22.15f}\n")
            for iw in range(p.Wr.shape[1]):
                fh.write(f"{f + 1:6d}{iw + 1:6d}{p.Wr[f, iw]:22.15f}{0.0:22.15f}\n")
            fh.write("\n")


def read_sym_info(path: str) -> tuple[np.ndarray, np.ndarray]:
    lines = [l for l in _data_lines(path)]
    P, ncopy = (int(t) for t in lines[1].split()[:2])
    if ncopy != 1:
        raise ValueError("Ncopy /= 1 (main.f90:93)")
    chi = np.empty(P, dtype=np.complex128)
    for g in range(P):
        tok = lines[3 + g].split(None, 1)
        if int(tok[0]) != g + 1:
            raise ValueError("i1/=mpidx (sym_info.txt)")
        re, im = tok[1].strip().strip("()").split(",")
        chi[g] = complex(float(re), float(im))
    rows = [l.split() for l in lines[4 + P:] if l.strip()]
    N = len(rows) // P
    w_map = np.empty((P, N), dtype=np.int32)
    for k, r in enumerate(rows):
        g, ic, i, w = (int(t) for t in r[:4])
        if g != k // N + 1 or ic != 1 or i != k % N + 1 or not 1 <= w <= N:
            raise ValueError("W_map wrong (sym_info.txt)")
        w_map[g - 1, i - 1] = w - 1
    for g in range(P):
        if sorted(w_map[g]) != list(range(N)):
            raise ValueError("W_map wrong 2 (sym_info.txt)")
    return w_map, chi


def read_orbitalidx(path: str, N: int) -> tuple[np.ndarray, int]:
    lines = _data_lines(path)
    nsl = int(lines[1].split()[1])
    orb = np.empty((N, N), dtype=np.int32)
    for k, l in enumerate(lines[5:5 + N * N]):
        i, j, o = (int(t) for t in l.split()[:3])
        if i != k // N or j != k % N or not 0 <= o < nsl:
            raise ValueError("error, orbitalidx")
        orb[i, j] = o
    return orb, nsl


def read_zqp_orbital(path: str, n_slater: int) -> np.ndarray:
    lines = _data_lines(path)
    f = np.empty(n_slater)
    for k, l in enumerate(lines[5:5 + n_slater]):
        t = l.split()
        if int(t[0]) != k or abs(float(t[2])) > 1e-10:
            raise ValueError("error in zqp_orbital")
        f[k] = float(t[1])
    return f


def read_w(path: str, alphac: int, alphar: int, N: int) -> tuple:
    lines = _data_lines(path)
    k = 1
    Wc = np.zeros((alphac, N), np.complex128); bc = np.zeros(alphac, np.complex128)
    Wr = np.zeros((alphar, N)); br = np.zeros(alphar)
    for f in range(alphac):
        t = lines[k].split(); k += 1
        bc[f] = complex(float(t[3]), float(t[4]))
        for iw in range(N):
            t = lines[k].split(); k += 1
            Wc[f, iw] = complex(float(t[2]), float(t[3]))
        k += 1
    for f in range(alphar):
        t = lines[k].split(); k += 1
        br[f] = float(t[3])
        for iw in range(N):
            t = lines[k].split(); k += 1
            Wr[f, iw] = float(t[2])
        k += 1
    return Wc, bc, Wr, br


def read_rbm_input(path: str) -> dict:
    lines = _data_lines(path)
    L, ns, sp = lines[1].split()[:3]
    ac, ar, nstep, nav, dt = lines[3].split()[:5]
    s0, ss, sc = lines[5].split()[:3]
    J1, J2 = lines[7].split()[:2]
    return dict(L=int(L), nsample=int(ns), spin_parity=int(sp), alphac=int(ac),
                alphar=int(ar), nstep=int(nstep), nav=int(nav),
                delta_tau=float(dt.replace("d", "e").replace("D", "e")),
                s_factor0=float(s0.replace("d", "e")), s_factor_scale=float(ss.replace("d", "e")),
                s_cut=float(sc.replace("d", "e")), J1=float(J1.replace("d", "e")),
                J2=float(J2.replace("d", "e")))


def write_run_dir(path: str, spec: ModelSpec, p: Params) -> None:
    os.makedirs(path, exist_ok=True)
    write_rbm_input(os.path.join(path, "RBM.input"), spec)
    write_sym_info(os.path.join(path, "sym_info.txt"), spec.w_map, spec.chi)
    write_orbitalidx(os.path.join(path, "orbitalidx.def"), spec.orb, spec.n_slater)
    write_zqp_orbital(os.path.join(path, "zqp_orbital_initial.dat"), p.f)
    write_w(os.path.join(path, "W.input"), p)


def read_run_dir(path: str) -> tuple[ModelSpec, Params]:
    d = read_rbm_input(os.path.join(path, "RBM.input"))
    N = d["L"] ** 2
    w_map, chi = read_sym_info(os.path.join(path, "sym_info.txt"))
    orb, nsl = read_orbitalidx(os.path.join(path, "orbitalidx.def"), N)
    spec = ModelSpec(L=d["L"], alphac=d["alphac"], alphar=d["alphar"], J1=d["J1"],
                     J2=d["J2"], spin_parity=d["spin_parity"], w_map=w_map, chi=chi,
                     orb=orb, n_slater=nsl, nsample=d["nsample"], nstep=d["nstep"],
                     nav=d["nav"], delta_tau=d["delta_tau"], s_factor0=d["s_factor0"],
                     s_factor_scale=d["s_factor_scale"], s_cut=d["s_cut"])
    f = renormalise_f(read_zqp_orbital(os.path.join(path, "zqp_orbital_initial.dat"), nsl))
    wpath = os.path.join(path, "W.input")
    if os.path.exists(wpath):
        Wc, bc, Wr, br = read_w(wpath, spec.alphac, spec.alphar, N)
    else:
        p0 = random_params(spec)
        Wc, bc, Wr, br = p0.Wc, p0.bc, p0.Wr, p0.br
    return spec, Params(Wc=Wc, bc=bc, Wr=Wr, br=br, f=f)
