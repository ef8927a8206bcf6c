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
