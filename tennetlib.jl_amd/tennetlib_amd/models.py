"""MPO builders and product states for the benchmark models.

The reference builds MPOs with ITensors' OpSum (e.g. test/test_MPS_DMRG.jl:26-43);
here the XXZ/Heisenberg chain is written directly as its w=5 finite-state
machine: W[wl, s', s, wr] with the chain starting in state 4 and ending in
state 0, so the first tensor is row 4 (1, 2, 2, 5) and the last column 0
(5, 2, 2, 1).  Basis |up> = 0, |down> = 1.
"""

from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np


def xxz_mpo(N: int, J: float = 1.0, Delta: float = 1.0, h: Optional[Sequence[float]] = None) -> List[np.ndarray]:
    """H = sum_i J/2 (S+_i S-_{i+1} + h.c.) + J Delta Sz_i Sz_{i+1} + sum_i h_i Sz_i."""
    I = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    h = np.zeros(N) if h is None else np.asarray(h, dtype=np.float64)
    out = []
    for k in range(N):
        W = np.zeros((5, 2, 2, 5))
        W[0, :, :, 0] = I
        W[1, :, :, 0] = Sp
        W[2, :, :, 0] = Sm
        W[3, :, :, 0] = Sz
        W[4, :, :, 0] = h[k] * Sz
        W[4, :, :, 1] = 0.5 * J * Sm
        W[4, :, :, 2] = 0.5 * J * Sp
        W[4, :, :, 3] = J * Delta * Sz
        W[4, :, :, 4] = I
        if k == 0:
            W = W[4:5]
        if k == N - 1:
            W = W[..., 0:1]
        out.append(np.ascontiguousarray(W))
    return out


def heisenberg_mpo(N: int, J: float = 1.0) -> List[np.ndarray]:
    return xxz_mpo(N, J=J, Delta=1.0)


def product_state(vectors: Sequence[np.ndarray], dtype=np.float64) -> List[np.ndarray]:
    return [np.asarray(v, dtype=dtype).reshape(1, -1, 1).copy() for v in vectors]


def neel_state(N: int, d: int = 2, dtype=np.float64) -> List[np.ndarray]:
    vs = []
    for k in range(N):
        v = np.zeros(d)
        v[k % 2] = 1.0
        vs.append(v)
    return product_state(vs, dtype)
