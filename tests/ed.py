"""Exact diagonalisation of the J1-J2 Heisenberg model in the Sz=0 sector.

Independent of the oracle: builds the Hamiltonian in Pauli units
(sigma.sigma = 4 S.S) from the bond list alone, in either the plain basis or
the sublattice-rotated (Marshall) gauge used by the reference's local energy
(calc_phys_quantity.f90:332-375: off-diagonal 2*J*sgn with sgn=-1 on J1 bonds).
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def sz0_states(N: int) -> np.ndarray:
    """All configurations with N/2 up spins, as int8 arrays of +-1 [n, N]."""
    out = []
    for ups in itertools.combinations(range(N), N // 2):
        x = -np.ones(N, dtype=np.int8)
        x[list(ups)] = 1
        out.append(x)
    return np.array(out)


def hamiltonian(L: int, J1: float, J2: float, gauge: bool):
    N = L * L
    states = sz0_states(N)
    weights = 1 << np.arange(N, dtype=np.int64)
    keys = ((states > 0).astype(np.int64) * weights).sum(1)
    index = {int(k): i for i, k in enumerate(keys)}
    bonds = []
    for i in range(N):
        ix, iy = divmod(i, L)
        bonds.append((i, ((ix + 1) % L) * L + iy, J1, -1.0))
        bonds.append((i, ix * L + (iy + 1) % L, J1, -1.0))
        bonds.append((i, ((ix + 1) % L) * L + (iy + 1) % L, J2, 1.0))
        bonds.append((i, ((ix + 1) % L) * L + (iy - 1) % L, J2, 1.0))
    rows, cols, vals = [], [], []
    diag = np.zeros(len(states))
    for a, b, J, sg in bonds:
        diag += J * states[:, a] * states[:, b]
        anti = np.nonzero(states[:, a] != states[:, b])[0]
        for s in anti:
            k = int(keys[s]) ^ (1 << a) ^ (1 << b)
            rows.append(index[k]); cols.append(s)
            vals.append(2.0 * J * (sg if gauge else 1.0))
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states),) * 2)
    H = H + sp.diags(diag)
    return states, H


def ground_energy(L: int, J1: float, J2: float) -> float:
    _, H = hamiltonian(L, J1, J2, gauge=False)
    return float(sla.eigsh(H, k=1, which="SA")[0][0])


def pair_operators(L: int, gauge: bool):
    """For every pair i1 <= i2 (i2 outer, i1 inner) on the Sz=0 states:
    the diagonal of sigma^z_i1 sigma^z_i2 and the sparse exchange operator
    (sigma^x sigma^x + sigma^y sigma^y) / 2 of the pair, i.e. |x> -> |x'> for
    antiparallel spins (times s_i1 s_i2, s = -1 on sublattice B, in the
    gauge of the Hamiltonian above) and the identity for i1 == i2."""
    N = L * L
    states = sz0_states(N)
    weights = 1 << np.arange(N, dtype=np.int64)
    keys = ((states > 0).astype(np.int64) * weights).sum(1)
    index = {int(k): i for i, k in enumerate(keys)}
    sub = np.array([((i // L) + (i % L)) & 1 for i in range(N)])
    out = []
    n = len(states)
    for i2 in range(N):
        for i1 in range(i2 + 1):
            zz = (states[:, i1] * states[:, i2]).astype(np.float64)
            if i1 == i2:
                ex = sp.identity(n, format="csr")
            else:
                sg = -1.0 if (gauge and sub[i1] != sub[i2]) else 1.0
                anti = np.nonzero(states[:, i1] != states[:, i2])[0]
                rows = [index[int(keys[s]) ^ (1 << i1) ^ (1 << i2)] for s in anti]
                ex = sp.csr_matrix((np.full(len(anti), sg), (rows, anti)), shape=(n, n))
            out.append((i1, i2, zz, ex))
    return states, out
