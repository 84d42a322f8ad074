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


def hubbard_site_ops():
    """Spinful-fermion site in the basis |0>, |up>, |dn>, |updn> = c+_up c+_dn |0>:
    (I, F, c_up, c_dn) with F = (-1)^n the Jordan-Wigner parity; Op[s', s]."""
    I = np.eye(4)
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    cup = np.zeros((4, 4))
    cup[0, 1] = 1.0   # |up> -> |0>
    cup[2, 3] = 1.0   # |updn> -> |dn>
    cdn = np.zeros((4, 4))
    cdn[0, 2] = 1.0   # |dn> -> |0>
    cdn[1, 3] = -1.0  # c_dn c+_up c+_dn |0> = -c+_up |0>
    return I, F, cup, cdn


def hubbard_mpo(N: int, t: float = 1.0, U: float = 4.0) -> List[np.ndarray]:
    """1D Fermi-Hubbard chain H = -t sum_{i,s} (c+_{i,s} c_{i+1,s} + h.c.) +
    U sum_i n_{i,up} n_{i,dn}, open boundaries, as a w = 6 MPO with
    Jordan-Wigner strings (the config of the QN block-sparse benchmark:
    N = 64, t = 1, U = 4).  Channel 5 is "nothing placed yet", 0 "done";
    channels 1-4 carry an open hopping term: c+_s F or c_s F was placed at
    site i (the string F is attached on the operator's own site, so the
    partner at i+1 needs none), the partner c_s or c+_s closes it."""
    I, F, cup, cdn = hubbard_site_ops()
    nup, ndn = cup.T @ cup, cdn.T @ cdn
    out = []
    for k in range(N):
        W = np.zeros((6, 4, 4, 6))
        W[0, :, :, 0] = I
        W[1, :, :, 0] = cup
        W[2, :, :, 0] = cup.T
        W[3, :, :, 0] = cdn
        W[4, :, :, 0] = cdn.T
        W[5, :, :, 0] = U * (nup @ ndn)
        W[5, :, :, 1] = -t * (cup.T @ F)
        W[5, :, :, 2] = t * (cup @ F)
        W[5, :, :, 3] = -t * (cdn.T @ F)
        W[5, :, :, 4] = t * (cdn @ F)
        W[5, :, :, 5] = I
        if k == 0:
            W = W[5:6]
        if k == N - 1:
            W = W[..., 0:1]
        out.append(np.ascontiguousarray(W))
    return out


def spin_half_qns():
    """QNs of |up>, |down>: (2 Sz,) = (1,), (-1,)."""
    return [(1,), (-1,)]


def hubbard_qns():
    """QNs of |0>, |up>, |dn>, |updn>: (N_up, N_dn)."""
    return [(0, 0), (1, 0), (0, 1), (1, 1)]


def hubbard_occupation_state(occ: Sequence[int]) -> List[np.ndarray]:
    """Product state with site k in basis state occ[k] (0 empty, 1 up, 2 dn, 3 both)."""
    return product_state([np.eye(4)[o] for o in occ])


def product_state(vectors: Sequence[np.ndarray], dtype=np.float64) -> List[np.ndarray]:
    return [np.asarray(v, dtype=dtype).reshape(1, -1, 1).copy() for v in vectors]


def neel_state(N: int, d: int = 2, dtype=np.float64) -> List[np.ndarray]:
    vs = []
    for k in range(N):
        v = np.zeros(d)
        v[k % 2] = 1.0
        vs.append(v)
    return product_state(vs, dtype)


def spin_coupling_mpo(N: int, couplings, h: Optional[Sequence[float]] = None, compress: bool = True,
                      tol: float = 1e-13, conserve_sz: bool = False) -> List[np.ndarray]:
    """H = sum_{(i,j)} J_ij S_i . S_j + sum_i h_i Sz_i for arbitrary pairs
    (0-based sites, i != j), as an MPO.

    Finite-state machine with one channel per (operator, source site) for
    every site that still has partners beyond the bond: states are
    0 = done, 1..3|A_k| = (S+, S-, Sz) emitted at a source i <= k, last =
    start; W = 2 + 3|A_k|.  This plays the role of the reference's OpSum ->
    MPO input (e.g. its J1-J2 cylinder models).  With `compress`, each bond
    is then reduced by an SVD sweep pair (relative tolerance `tol` on the
    singular values of the bond matrices), which keeps H exact to
    ~tol * |H| and gives the minimal bond dimension of the sum.  With
    `conserve_sz` the compression keeps every bond channel of definite S_z
    flux (each bond matrix is block-diagonal in the flux, and the QR and SVD
    run block by block), so the MPO can drive an S_z-conserving state
    (`StateEnvs(..., qns=[spin_half_qns()] * N)`); the bond dimension is the
    same, since the minimal rank of a block-diagonal matrix is the sum of its
    blocks' ranks."""
    I = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    J = {}
    for (i, j), v in dict(couplings).items():
        if i == j:
            raise ValueError("self coupling")
        a, b = (i, j) if i < j else (j, i)
        J[(a, b)] = J.get((a, b), 0.0) + float(v)
    h = np.zeros(N) if h is None else np.asarray(h, dtype=np.float64)
    partners = [[] for _ in range(N)]  # partners j > i
    for (a, b), v in J.items():
        if v != 0.0:
            partners[a].append(b)
    # active sources across bond k (between k and k+1): i <= k with a partner > k
    active = [[i for i in range(k + 1) if any(j > k for j in partners[i])] for k in range(N)]
    # emitted operator at the source and the partner's operator with its weight
    chans = [(Sp, Sm, 0.5), (Sm, Sp, 0.5), (Sz, Sz, 1.0)]
    Ws = []
    for k in range(N):
        left = active[k - 1] if k > 0 else []
        right = active[k] if k < N - 1 else []
        wl, wr = 2 + 3 * len(left), 2 + 3 * len(right)
        W = np.zeros((wl, 2, 2, wr))
        sl, sr = wl - 1, wr - 1  # start states
        W[0, :, :, 0] = I
        W[sl, :, :, sr] = I
        W[sl, :, :, 0] = h[k] * Sz
        ridx = {i: n for n, i in enumerate(right)}
        lidx = {i: n for n, i in enumerate(left)}
        if k in ridx:
            for c, (src, _, _) in enumerate(chans):
                W[sl, :, :, 1 + 3 * ridx[k] + c] = src
        for i, n in lidx.items():
            for c, (_, dst, wt) in enumerate(chans):
                if i in ridx:
                    W[1 + 3 * n + c, :, :, 1 + 3 * ridx[i] + c] = I
                v = J.get((i, k), 0.0)
                if v != 0.0:
                    W[1 + 3 * n + c, :, :, 0] = wt * v * dst
        if k == 0:
            W = W[sl:sl + 1]
        if k == N - 1:
            W = W[..., 0:1]
        Ws.append(np.ascontiguousarray(W))
    if compress and N > 1:
        Ws = _compress_mpo_flux(Ws, tol, [1, -1]) if conserve_sz else _compress_mpo(Ws, tol)
    return Ws


def _mpo_channel_flux(Ws: List[np.ndarray], q: Sequence[int]) -> List[np.ndarray]:
    """Flux of every bond channel of an MPO whose entries W[a, s, t, b] change
    the flux by q[s] - q[t]: flux[k + 1][b] = flux[k][a] + q[s] - q[t] for
    every nonzero entry (flux of the left edge 0).  Raises if a channel would
    carry two fluxes."""
    fl = [np.zeros(Ws[0].shape[0], dtype=np.int64)]
    for k, W in enumerate(Ws):
        nxt = {}
        for a, s, t, b in zip(*np.nonzero(W)):
            f = int(fl[-1][a]) + q[s] - q[t]
            if nxt.setdefault(int(b), f) != f:
                raise ValueError(f"MPO channel {b} of bond {k + 1} has no definite flux")
        fl.append(np.array([nxt.get(b, 0) for b in range(W.shape[3])], dtype=np.int64))
    return fl


def _compress_mpo_flux(Ws: List[np.ndarray], tol: float, q: Sequence[int]) -> List[np.ndarray]:
    """_compress_mpo block by block in the channel flux: the bond matrices are
    block-diagonal once rows and columns are grouped by flux, so the QR of a
    column block and the SVD of a row block leave every new channel with one
    flux; singular values are cut at tol times the bond's largest."""
    Ws = [w.copy() for w in Ws]
    N = len(Ws)
    q = list(q)
    fl = _mpo_channel_flux(Ws, q)
    dq = np.subtract.outer(np.array(q), np.array(q))  # q[s] - q[t]
    for k in range(N - 1):  # left-orthogonalise with a QR per flux block of the right channels
        wl, d1, d2, wr = Ws[k].shape
        rowf = (fl[k][:, None, None] + dq[None, :, :]).reshape(-1)
        M = Ws[k].reshape(wl * d1 * d2, wr)
        Qs, Rs, nf = [], [], []
        for f in np.unique(fl[k + 1]):
            cols = np.nonzero(fl[k + 1] == f)[0]
            rows = np.nonzero(rowf == f)[0]
            Qb, Rb = np.linalg.qr(M[np.ix_(rows, cols)])
            Qf = np.zeros((M.shape[0], Qb.shape[1]))
            Qf[rows] = Qb
            Rf = np.zeros((Qb.shape[1], wr))
            Rf[:, cols] = Rb
            Qs.append(Qf)
            Rs.append(Rf)
            nf += [int(f)] * Qb.shape[1]
        Q, R = np.hstack(Qs), np.vstack(Rs)
        Ws[k] = Q.reshape(wl, d1, d2, -1)
        Ws[k + 1] = np.tensordot(R, Ws[k + 1], axes=([1], [0]))
        fl[k + 1] = np.array(nf, dtype=np.int64)
    for k in range(N - 1, 0, -1):  # right-to-left truncated SVD per flux block of the left channels
        wl, d1, d2, wr = Ws[k].shape
        colf = (fl[k + 1][None, None, :] - dq[:, :, None]).reshape(-1)
        M = Ws[k].reshape(wl, d1 * d2 * wr)
        blocks = []
        for f in np.unique(fl[k]):
            rows = np.nonzero(fl[k] == f)[0]
            cols = np.nonzero(colf == f)[0]
            U, sv, Vt = np.linalg.svd(M[np.ix_(rows, cols)], full_matrices=False)
            blocks.append((int(f), rows, cols, U, sv, Vt))
        smax = max(b[4][0] for b in blocks if b[4].size)
        Vs, Us, nf = [], [], []
        for f, rows, cols, U, sv, Vt in blocks:
            keep = int(np.sum(sv > tol * smax))
            if keep == 0:
                continue
            Vf = np.zeros((keep, M.shape[1]))
            Vf[:, cols] = Vt[:keep]
            Uf = np.zeros((wl, keep))
            Uf[rows] = U[:, :keep] * sv[:keep]
            Vs.append(Vf)
            Us.append(Uf)
            nf += [f] * keep
        V, Usv = np.vstack(Vs), np.hstack(Us)
        Ws[k] = V.reshape(-1, d1, d2, wr)
        Ws[k - 1] = np.tensordot(Ws[k - 1], Usv, axes=([3], [0]))
        fl[k] = np.array(nf, dtype=np.int64)
    return [np.ascontiguousarray(w) for w in Ws]


def _compress_mpo(Ws: List[np.ndarray], tol: float) -> List[np.ndarray]:
    """Left-to-right QR then right-to-left truncated SVD of the MPO bonds."""
    Ws = [w.copy() for w in Ws]
    N = len(Ws)
    for k in range(N - 1):  # left-orthogonalise with QR
        wl, d1, d2, wr = Ws[k].shape
        Q, R = np.linalg.qr(Ws[k].reshape(wl * d1 * d2, wr))
        Ws[k] = Q.reshape(wl, d1, d2, -1)
        Ws[k + 1] = np.tensordot(R, Ws[k + 1], axes=([1], [0]))
    for k in range(N - 1, 0, -1):  # right-to-left truncated SVD
        wl, d1, d2, wr = Ws[k].shape
        U, s, Vt = np.linalg.svd(Ws[k].reshape(wl, d1 * d2 * wr), full_matrices=False)
        keep = max(1, int(np.sum(s > tol * s[0])))
        Ws[k] = Vt[:keep].reshape(keep, d1, d2, wr)
        Ws[k - 1] = np.tensordot(Ws[k - 1], U[:, :keep] * s[:keep], axes=([3], [0]))
    return [np.ascontiguousarray(w) for w in Ws]


def cylinder_couplings(Ly: int, Lx: int, J1: float = 1.0, J2: float = 0.0, snake: bool = True) -> dict:
    """J1-J2 couplings on an Ly x Lx square-lattice cylinder, periodic in y
    (Ly >= 3), open in x; sites numbered column by column, reversing the
    y order in every other column when `snake`."""
    def idx(x, y):
        y %= Ly
        return x * Ly + (y if (not snake or x % 2 == 0) else Ly - 1 - y)
    J = {}

    def add(a, b, v):
        key = (min(a, b), max(a, b))
        J[key] = J.get(key, 0.0) + v
    for x in range(Lx):
        for y in range(Ly):
            add(idx(x, y), idx(x, y + 1), J1)  # vertical (periodic)
            if x + 1 < Lx:
                add(idx(x, y), idx(x + 1, y), J1)  # horizontal
                if J2 != 0.0:
                    add(idx(x, y), idx(x + 1, y + 1), J2)
                    add(idx(x, y), idx(x + 1, y - 1), J2)
    return J


def j1j2_cylinder_mpo(Ly: int, Lx: int, J1: float = 1.0, J2: float = 0.5, snake: bool = True,
                      compress: bool = True, conserve_sz: bool = False) -> List[np.ndarray]:
    """The J1-J2 Heisenberg cylinder of the wide-MPO benchmark (Ly = 6, Lx = 12,
    J2 = 0.5 there); `conserve_sz` keeps S_z flux per bond channel through the
    compression."""
    return spin_coupling_mpo(Ly * Lx, cylinder_couplings(Ly, Lx, J1, J2, snake), compress=compress,
                             conserve_sz=conserve_sz)
