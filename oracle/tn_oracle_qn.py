"""FP64 oracle for U(1) / U(1)xU(1) quantum-number (QN) conserving two-site
DMRG -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may import
this module, as the checker; no product path calls it.

It restates the block-sparse behaviour of the reference's QN-conserving runs
(test/test_MPS_DMRG.jl:10,29 `conserve_qns = qn`, default true), which the
reference gets from ITensors/NDTensors BlockSparse tensors (absent here):

* every index carries a QN per basis element (a tuple of ints: (2 Sz,) for
  spin-1/2, (N_up, N_dn) for the Hubbard site); an MPS tensor A[a, s, b] may
  be nonzero only where q(a) + q(s) = q(b) (link b carries the QN of the
  sites to its left);
* the products, environments and Lanczos vectors are those of the dense
  oracle (tn_oracle): a state with definite QN stays exactly block-sparse
  under a QN-conserving MPO, zeros stay exact zeros;
* replacebond! factorises block by block: the density matrix of the kept
  side is block diagonal in the QN of the combined (link, site) index; every
  block is diagonalised, all weights go through NDTensors truncate!!
  (truncate_weights: maxdim, relative cutoff, mindim) and each block keeps
  its weights above docut (NDTensors _truncated_blockdim), so degenerate
  multiplets split across blocks are kept or dropped together and exactly
  zero weights are dropped even without truncation; the new link lists the
  kept sectors in ascending QN order, each with its weights in descending
  order (an ordering choice of this oracle; energies, weights and sector
  dimensions do not depend on it).

Pins (tests/test_oracle_qn.py): the block spectra are the dense spectrum;
the free-fermion (U = 0) Hubbard ground energy at half filling; exact
diagonalisation of small chains in the target sector; the sector dimensions
of exact (untruncated) states equal the binomial counts of the basis.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from oracle import tn_oracle as O

QN = Tuple[int, ...]


def spin_half_qns() -> List[QN]:
    """|up>, |down> of tn_oracle.spin_half_ops: 2 Sz = +1, -1."""
    return [(1,), (-1,)]


def hubbard_qns() -> List[QN]:
    """|0>, |up>, |dn>, |updn> of tn_oracle.hubbard_ops: (N_up, N_dn)."""
    return [(0, 0), (1, 0), (0, 1), (1, 1)]


def qadd(a: QN, b: QN) -> QN:
    return tuple(x + y for x, y in zip(a, b))


def qsub(a: QN, b: QN) -> QN:
    return tuple(x - y for x, y in zip(a, b))


def link_qns_of_product(A: List[np.ndarray], sq: List[List[QN]]) -> List[List[QN]]:
    """Link QNs (bonds 0..N) of an MPS whose every link element has a definite
    QN, read off its nonzero entries from the left (q_0 = 0)."""
    nq = len(sq[0][0])
    links = [[tuple([0] * nq)]]
    for k, a in enumerate(A):
        ql = links[-1]
        qr: List[Optional[QN]] = [None] * a.shape[2]
        for (i, s, j) in zip(*np.nonzero(a)):
            q = qadd(ql[i], sq[k][s])
            if qr[j] is None:
                qr[j] = q
            elif qr[j] != q:
                raise ValueError(f"site {k + 1}: link element {j} has no definite QN")
        links.append([q if q is not None else tuple([0] * nq) for q in qr])
    return links


def check_flux(A: List[np.ndarray], sq: List[List[QN]], links: List[List[QN]], tol=0.0) -> bool:
    """True when every entry outside the QN-allowed blocks is (at most tol) zero."""
    for k, a in enumerate(A):
        ql, qr = links[k], links[k + 1]
        for i in range(a.shape[0]):
            for s in range(a.shape[1]):
                q = qadd(ql[i], sq[k][s])
                bad = [j for j in range(a.shape[2]) if qr[j] != q]
                if bad and np.abs(a[i, s, bad]).max() > tol:
                    return False
    return True


@dataclass
class BlockSpectrum:
    eigs: np.ndarray          # kept weights, descending (all blocks)
    truncerr: float
    sectors: Dict[QN, int]    # kept dimension per sector of the new link


def block_truncate(blocks: Dict[QN, np.ndarray], maxdim: int, mindim: int, cutoff: float):
    """NDTensors block truncation: truncate!! on all weights (descending),
    then each block (weights descending) keeps its leading weights > docut
    (_truncated_blockdim); blocks keeping nothing are dropped.  Returns
    ({sector: kept}, truncerr)."""
    allP = np.sort(np.concatenate([np.asarray(w) for w in blocks.values()]))[::-1]
    tr = O.truncate_weights(allP, maxdim, mindim, cutoff)
    kept = {}
    for q in sorted(blocks):
        w = blocks[q]
        k = 0
        while k < len(w) and w[k] > tr.docut:
            k += 1
        if k:
            kept[q] = k
    return kept, tr.truncerr


def block_factorize(phi: np.ndarray, ql: List[QN], qs1: List[QN], qs2: List[QN], qr: List[QN],
                    maxdim: int, mindim: int, cutoff: float, ortho: str, normalize: bool = True):
    """replacebond! of a QN-conserving two-site tensor phi (chil, d1, d2, chir)
    by block density matrices (see the module docstring).  Returns
    (A_b, A_{b+1}, new link QNs, BlockSpectrum)."""
    a, s, t, c = phi.shape
    M = phi.reshape(a * s, t * c)
    if ortho == "left":
        q_idx = [qadd(ql[i], qs1[j]) for i in range(a) for j in range(s)]
        G = M
    elif ortho == "right":
        q_idx = [qsub(qr[k], qs2[j]) for j in range(t) for k in range(c)]
        G = M.T
    else:
        raise ValueError("ortho must be 'left' or 'right'")
    sectors = sorted(set(q_idx))
    rows = {q: [i for i, x in enumerate(q_idx) if x == q] for q in sectors}
    blocks = {}
    for q in sectors:
        Gq = G[rows[q]]
        rho = Gq @ Gq.conj().T
        w, V = np.linalg.eigh(0.5 * (rho + rho.conj().T))
        blocks[q] = (w[::-1].copy(), V[:, ::-1].copy())
    kept, truncerr = block_truncate({q: blocks[q][0] for q in sectors}, maxdim, mindim, cutoff)
    newq: List[QN] = []
    U = np.zeros((G.shape[0], sum(kept.values())), dtype=phi.dtype)
    col = 0
    eigs = []
    for q in sectors:
        if q not in kept:
            continue
        k = kept[q]
        U[rows[q], col:col + k] = blocks[q][1][:, :k]
        eigs.extend(blocks[q][0][:k])
        newq.extend([q] * k)
        col += k
    n = col
    if ortho == "left":
        R = U.conj().T @ M
        if normalize:
            R = R / np.linalg.norm(R)
        Ab, Ab1 = U.reshape(a, s, n), R.reshape(n, t, c)
    else:
        Lf = M @ U.conj()
        if normalize:
            Lf = Lf / np.linalg.norm(Lf)
        Ab, Ab1 = Lf.reshape(a, s, n), U.T.reshape(n, t, c)
    P = np.array(sorted(eigs, reverse=True))
    return Ab, Ab1, newq, BlockSpectrum(P, truncerr, dict(kept))


class QNState:
    """O.StateEnvs plus the QNs of the site and link indices."""

    def __init__(self, psi: List[np.ndarray], Ws: List[np.ndarray], site_qns: List[List[QN]]):
        self.env = O.StateEnvs(psi, Ws)
        self.sq = [list(q) for q in site_qns]
        self.links = link_qns_of_product(self.env.psi, self.sq)

    @property
    def psi(self):
        return self.env.psi

    def sectors(self, bond: int) -> Dict[QN, int]:
        out: Dict[QN, int] = {}
        for q in self.links[bond]:
            out[q] = out.get(q, 0) + 1
        return out


def _ortho_first(st: QNState):
    """Right-orthogonalize by block-preserving QR sweeps (exact, no truncation):
    each site's (chil) x (d chir) matrix is decomposed per left-link sector."""
    env = st.env
    if env.ortho_center == 1:
        return
    A = env.psi
    N = len(A)
    for k in range(N - 1, 0, -1):
        a, d, c = A[k].shape
        M = A[k].reshape(a, d * c)
        ql = st.links[k]
        sectors = sorted(set(ql))
        Rfull = np.zeros((a, a), dtype=M.dtype)
        Bnew = np.zeros_like(M)
        newq: List[QN] = []
        col = 0
        # per sector: M_q (rows with that link QN) = R_q Q_q, Q_q orthonormal rows
        blocks = []
        for q in sectors:
            r = [i for i, x in enumerate(ql) if x == q]
            Mq = M[r]
            Qt, Rt = np.linalg.qr(Mq.T)   # Mq^T = Qt Rt -> Mq = Rt^T Qt^T
            blocks.append((q, r, Qt.T, Rt.T))
        n = sum(b[2].shape[0] for b in blocks)
        Bnew = np.zeros((n, d * c), dtype=M.dtype)
        Rm = np.zeros((a, n), dtype=M.dtype)
        for q, r, Qq, Rq in blocks:
            m = Qq.shape[0]
            Bnew[col:col + m] = Qq
            Rm[np.ix_(r, range(col, col + m))] = Rq
            newq.extend([q] * m)
            col += m
        A[k] = Bnew.reshape(n, d, c)
        A[k - 1] = np.tensordot(A[k - 1], Rm, axes=([2], [0]))
        st.links[k] = newq
    env.ortho_center = 1
    env.PH = env.PH.fresh()


def update_two_site(st: QNState, pos: int, ortho: str, *, maxdim=2 ** 62, mindim=1,
                    cutoff=O.FLOAT64_THRESHOLD, krylovdim=5, maxiter=2, tol=1e-14, normalize=True):
    """_update_two_site! (src/mps/update_site.jl:27-90) with the block
    factorisation; eig_solver only."""
    env = st.env
    A = env.psi
    env.PH.nsite = 2
    phi = np.tensordot(A[pos - 1], A[pos], axes=([2], [0]))
    env.PH.position(A, pos)
    energy, phi, _ = O.eigsolve_lanczos(env.PH.product, phi, krylovdim, maxiter, tol)
    if normalize:
        phi = phi / np.linalg.norm(phi)
    Ab, Ab1, newq, spec = block_factorize(phi, st.links[pos - 1], st.sq[pos - 1], st.sq[pos], st.links[pos + 1],
                                          maxdim, mindim, cutoff, ortho, normalize)
    A[pos - 1], A[pos] = Ab, Ab1
    st.links[pos] = newq
    env.ortho_center = pos + 1 if ortho == "left" else pos
    return energy, spec


def fullsweep(st: QNState, swdata: O.SweepData, **kw):
    """fullsweep! (src/mps/sweep.jl:94-196) with block factorisations."""
    _ortho_first(st)
    N = len(st.psi)
    maxtruncerr = 0.0
    swdata.sweepcount += 1
    lasteigs: List[np.ndarray] = [None] * (N - 1)
    energy = float("nan")
    for bond in range(1, N):
        energy, spec = update_two_site(st, bond, "left", **kw)
        maxtruncerr = max(spec.truncerr, maxtruncerr)
    for bond in range(N - 1, 0, -1):
        energy, spec = update_two_site(st, bond, "right", **kw)
        lasteigs[bond - 1] = spec.eigs
        maxtruncerr = max(spec.truncerr, maxtruncerr)
    swdata.maxchi.append(max(a.shape[2] for a in st.psi))
    swdata.energy.append(energy)
    mid = lasteigs[N // 2 - 1]
    swdata.entropy.append(O.entropy(mid / mid.sum()))
    swdata.maxtruncerr.append(maxtruncerr)
    swdata.lasteigs = lasteigs


def dmrg(st: QNState, maxdim: Sequence[int], cutoff: float, krylovdim=5, maxiter=2) -> O.SweepData:
    """dmrg! (src/mps/dmrg.jl:148-227) without noise, one sweep per maxdim entry."""
    sw = O.SweepData()
    for md in maxdim:
        fullsweep(st, sw, maxdim=md, cutoff=cutoff, krylovdim=krylovdim, maxiter=maxiter)
    return sw


def hubbard_free_fermion_energy(N: int, nup: int, ndn: int, t: float = 1.0) -> float:
    """Ground energy of the U = 0 open Hubbard chain: the lowest single-particle
    levels -2t cos(pi k / (N + 1)) filled by each spin species."""
    e = sorted(-2.0 * t * math.cos(math.pi * k / (N + 1)) for k in range(1, N + 1))
    return float(sum(e[:nup]) + sum(e[:ndn]))
