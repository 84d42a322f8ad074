"""FP64 CPU oracle for the effective-Hamiltonian hot path (numpy / scipy).

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP path is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only as the checker or the timed CPU
baseline -- never as the thing measured or shipped.  The product path
(``tennetlib.jl_amd``) must never route through this file.

What it restates (reference = TenNetLib.jl 0.2.2, files cited per function):

* MPO builders for the benchmark models (the reference builds them with
  ITensors' OpSum; here they are written as finite-state machines).
* Environment updates ``makeL!``/``makeR!`` and the two-/one-site projected
  products of ``ProjMPO`` (upstream ITensorMPS; in-tree mirror
  ``src/mps/projmps2.jl:77-173``).
* ``eig_solver`` (``src/base/solver.jl:23-43``): KrylovKit ``eigsolve`` with
  Lanczos, full re-orthogonalisation and thick restart.
* ``exp_solver`` (``src/base/solver.jl:66-88``): KrylovKit ``exponentiate``.
* ``replacebond!``/``svd``/``truncate!!`` (upstream ITensorMPS/NDTensors,
  called from ``src/mps/update_site.jl:64-76``).
* ``_update_two_site!`` (``src/mps/update_site.jl:27-90``), ``fullsweep!``
  (``src/mps/sweep.jl:94-196``), ``dmrg!`` (``src/mps/dmrg.jl:148-227``) and
  ``tdvpsweep!`` (``src/mps/tdvp.jl:247-277``).

Index conventions (row-major, all contiguous):
  A_k[a, s, b]           MPS site tensor (chi_{k-1}, d, chi_k)
  W_k[wl, s', s, wr]     MPO tensor, s' is the bra (output) index
  L[a, w, a']            left environment: ket a, MPO bond w, bra a'
  R[b, w, b']            right environment: ket b, MPO bond w, bra b'
  psi2[a, s, t, b]       two-site wave function

Pinning of each function (what fixes it besides the restatement itself) is
listed in the tests: exact diagonalisation, P^dagger H P projections, closed
forms and brute force.  KrylovKit's restart bookkeeping and NDTensors' gauge
choices are restated from their published algorithms; where only an import of
those packages could pin a detail the test says "parity unpinned".
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np
import scipy.linalg as sla

FLOAT64_THRESHOLD = 1e-15  # src/base/global_variables.jl:10


# ---------------------------------------------------------------------------
# Seeded RNG: SplitMix64 -> uniform doubles -> Box-Muller normals.
# ---------------------------------------------------------------------------

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014) stream of 64-bit words."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) from the top 53 bits."""
        out = np.empty(n)
        for i in range(n):
            out[i] = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return out

    def normal(self, n: int) -> np.ndarray:
        """n standard normals by Box-Muller on pairs of uniforms."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]  # (0, 1]
        u2 = u[1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        out = np.empty(2 * m)
        out[0::2] = z[:m]
        out[1::2] = z[m:]
        return out[:n]


def splitmix_uniform_fast(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """Vectorised SplitMix64 uniforms in [lo, hi): word i is the (i+1)-th
    output of the stream seeded with ``seed`` (same values as SplitMix64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# Local operators and MPO builders.
# ---------------------------------------------------------------------------

def spin_half_ops():
    """Basis |up>=0, |down>=1; Op[s', s] = <s'|O|s>."""
    I = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    return I, Sz, Sp, Sm


def xxz_mpo(N: int, J: float = 1.0, Delta: float = 1.0,
            h: Optional[Sequence[float]] = None) -> List[np.ndarray]:
    """H = sum_i J/2 (S+_i S-_{i+1} + S-_i S+_{i+1}) + J*Delta Sz_i Sz_{i+1}
    + sum_i h_i Sz_i, open chain, as a w=5 finite-state-machine MPO.

    W[wl, s', s, wr]; the chain starts in state 4 and ends in state 0, so the
    first tensor is row 4 of the bulk W (shape (1,2,2,5)) and the last is
    column 0 (shape (5,2,2,1)).  Delta=1 is the Heisenberg chain of
    test/test_MPS_DMRG.jl:26-43."""
    I, Sz, Sp, Sm = spin_half_ops()
    if h is None:
        h = np.zeros(N)
    Ws = []
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
        Ws.append(np.ascontiguousarray(W))
    if N == 1:
        Ws[0] = np.ascontiguousarray(Ws[0])
    return Ws


def heisenberg_mpo(N: int, J: float = 1.0) -> List[np.ndarray]:
    return xxz_mpo(N, J=J, Delta=1.0)


def hubbard_ops():
    """Spinful fermion site, basis |0>, |up>, |dn>, |updn> (up created first:
    |updn> = c^+_up c^+_dn |0>).  Returns I, F (parity), c_up, c_dn, n_up, n_dn."""
    I = np.eye(4)
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    cup = np.zeros((4, 4))
    cdn = np.zeros((4, 4))
    # c_up: |up> -> |0>, |updn> -> |dn>
    cup[0, 1] = 1.0
    cup[2, 3] = 1.0
    # c_dn: |dn> -> |0>, |updn> = c+up c+dn|0> -> c_dn c+up c+dn |0> = -c+up|0>
    cdn[0, 2] = 1.0
    cdn[1, 3] = -1.0
    nup = cup.T @ cup
    ndn = cdn.T @ cdn
    return I, F, cup, cdn, nup, ndn


def hubbard_mpo(N: int, t: float = 1.0, U: float = 4.0) -> List[np.ndarray]:
    """1D Fermi-Hubbard H = -t sum_{i,sigma} (c+_{i s} c_{i+1 s} + h.c.)
    + U sum_i n_up n_dn with Jordan-Wigner strings, w=6 MPO."""
    I, F, cup, cdn, nup, ndn = hubbard_ops()
    cupd, cdnd = cup.T, cdn.T
    Ws = []
    for k in range(N):
        W = np.zeros((6, 4, 4, 6))
        W[0, :, :, 0] = I
        # second operator of a hopping pair (site i+1), closing the channel
        W[1, :, :, 0] = cup          # after c+_up F  at site i
        W[2, :, :, 0] = cupd         # after F c_up ... (h.c.)
        W[3, :, :, 0] = cdn
        W[4, :, :, 0] = cdnd
        W[5, :, :, 0] = U * nup @ ndn
        # opening operators at site i (JW string F attached on site i)
        W[5, :, :, 1] = -t * (cupd @ F)
        W[5, :, :, 2] = t * (cup @ F)
        W[5, :, :, 3] = -t * (cdnd @ F)
        W[5, :, :, 4] = t * (cdn @ F)
        W[5, :, :, 5] = I
        if k == 0:
            W = W[5:6]
        if k == N - 1:
            W = W[..., 0:1]
        Ws.append(np.ascontiguousarray(W))
    return Ws


def mpo_to_dense(Ws: List[np.ndarray]) -> np.ndarray:
    """Brute-force dense H (d^N x d^N) from an MPO, for small N."""
    T = Ws[0][0]  # (s', s, w)
    T = T.reshape(T.shape[0], T.shape[1], T.shape[2])
    for W in Ws[1:]:
        # T[(S'), (S), w] x W[w, s', s, w2]
        T = np.einsum("ABw,wxyz->AxByz", T, W)
        a, x, b, y, z = T.shape
        T = T.reshape(a * x, b * y, z)
    return T[:, :, 0]


# ---------------------------------------------------------------------------
# MPS helpers.
# ---------------------------------------------------------------------------

def product_mps(states: Sequence[np.ndarray]) -> List[np.ndarray]:
    return [np.asarray(v, dtype=np.float64).reshape(1, -1, 1).copy() for v in states]


def neel_mps(N: int, d: int = 2) -> List[np.ndarray]:
    vs = []
    for k in range(N):
        v = np.zeros(d)
        v[k % 2] = 1.0
        vs.append(v)
    return product_mps(vs)


def random_product_mps(N: int, d: int, seed: int) -> List[np.ndarray]:
    """Each site a normalised Gaussian d-vector from SplitMix64/Box-Muller."""
    rng = SplitMix64(seed)
    vs = []
    for _ in range(N):
        v = rng.normal(d)
        vs.append(v / np.linalg.norm(v))
    return product_mps(vs)


def mps_to_dense(A: List[np.ndarray]) -> np.ndarray:
    v = A[0]
    for B in A[1:]:
        v = np.tensordot(v, B, axes=([v.ndim - 1], [0]))
    return v.reshape(-1)


def right_orthogonalize(A: List[np.ndarray]) -> List[np.ndarray]:
    """Bring the MPS to orthocenter 1 (first site) by LQ sweeps from the
    right (ITensorMPS ``orthogonalize!(psi, 1)``, called at
    src/mps/sweep.jl:100-102).  Gauge: each Q has orthonormal rows."""
    A = [a.copy() for a in A]
    for k in range(len(A) - 1, 0, -1):
        a, s, b = A[k].shape
        M = A[k].reshape(a, s * b)
        Q, Rm = np.linalg.qr(M.conj().T)  # M^H = Q R  ->  M = R^H Q^H
        kk = Q.shape[1]
        A[k] = Q.conj().T.reshape(kk, s, b)
        A[k - 1] = np.tensordot(A[k - 1], Rm.conj().T, axes=([2], [0]))
    return A


# ---------------------------------------------------------------------------
# Environments and projected products (ProjMPO makeL!/makeR!/product).
# ---------------------------------------------------------------------------

def edge_env(dtype=np.float64) -> np.ndarray:
    return np.ones((1, 1, 1), dtype=dtype)


def update_left_env(L: np.ndarray, A: np.ndarray, W: np.ndarray) -> np.ndarray:
    """L_k = L_{k-1} * A_k * W_k * dag(A_k'): three pairwise contractions in
    the order of ProjMPO makeL! (mirror src/mps/projmps2.jl:84-89)."""
    X = np.tensordot(L, A, axes=([0], [0]))            # [w, a', s, b]
    Y = np.tensordot(X, W, axes=([0, 2], [0, 2]))       # [a', b, s', w']
    Ln = np.tensordot(Y, A.conj(), axes=([0, 2], [0, 1]))  # [b, w', b']
    return Ln


def update_right_env(R: np.ndarray, B: np.ndarray, W: np.ndarray) -> np.ndarray:
    """R_k = R_{k+1} * B_k * W_k * dag(B_k') (mirror src/mps/projmps2.jl:116-121)."""
    X = np.tensordot(B, R, axes=([2], [0]))            # [a, t, w2, b']
    Y = np.tensordot(X, W, axes=([1, 2], [2, 3]))       # [a, b', w, t']
    Rn = np.tensordot(Y, B.conj(), axes=([1, 3], [2, 1]))  # [a, w, a']
    return Rn


def matvec2(L, W1, W2, R, psi):
    """Two-site projected product Hv = (((v L) W1) W2) R, noprime'd
    (ProjMPO.product; mirror src/mps/projmps2.jl:159-173)."""
    T = np.tensordot(L, psi, axes=([0], [0]))           # [w, a', s, t, b]
    T = np.tensordot(T, W1, axes=([0, 2], [0, 2]))       # [a', t, b, s', w1]
    T = np.tensordot(T, W2, axes=([4, 1], [0, 2]))       # [a', b, s', t', w2]
    T = np.tensordot(T, R, axes=([1, 4], [0, 1]))        # [a', s', t', b']
    return T


def matvec1(L, W, R, psi):
    """One-site projected product (used by the TDVP reverse step,
    src/mps/update_site.jl:78-87)."""
    T = np.tensordot(L, psi, axes=([0], [0]))           # [w, a', s, b]
    T = np.tensordot(T, W, axes=([0, 2], [0, 2]))        # [a', b, s', w1]
    T = np.tensordot(T, R, axes=([1, 3], [0, 1]))        # [a', s', b']
    return T


def matvec0(L, R, C):
    """Zero-site projected product on the bond matrix C[a, b] (the backward
    step of a one-site update, src/mps/update_site.jl:176-183):
    out[a', b'] = sum L[a, w, a'] C[a, b] R[b, w, b']."""
    return np.einsum("awx,ab,bwy->xy", L, C, R)


def matvec_flops(chil: int, chir: int, d: int, w: int) -> float:
    """Algorithmic real flops of one two-site matvec (multiply by 4 for
    complex): 2 chil chir d^2 w (chil + chir) + 4 chil chir d^3 w^2."""
    return 2.0 * chil * chir * d * d * w * (chil + chir) + 4.0 * chil * chir * d ** 3 * w * w


def env_flops(chil: int, chir: int, d: int, w: int) -> float:
    """Algorithmic flops of one environment update:
    2 chil^2 chir d w + 2 chil chir d^2 w^2 + 2 chil chir^2 d w."""
    return 2.0 * chil * chil * chir * d * w + 2.0 * chil * chir * d * d * w * w + 2.0 * chil * chir * chir * d * w


class Envs:
    """ProjMPO-like cache of left/right environments with lpos/rpos
    semantics (mirror src/mps/projmps2.jl:77-149).

    LR[k] for k in 0..N+1: LR[0] and LR[N+1] are the trivial edges; after
    makeL!(k) LR[k] holds sites 1..k; after makeR!(k) LR[k] holds sites k..N.
    Sites are 1-based as in the reference."""

    def __init__(self, Ws: List[np.ndarray], dtype=np.float64, Ms=None, weight: float = 0.0):
        self.W = Ws
        self.N = len(Ws)
        self.LR: List[Optional[np.ndarray]] = [None] * (self.N + 2)
        self.lpos = 0
        self.rpos = self.N + 1
        self.nsite = 2
        self.dtype = dtype
        # penalty projectors (ProjMPO_MPS2, src/mps/projmpo_mps2.jl:94-101)
        self.Ms = Ms or []
        self.weight = weight
        self.m: List[np.ndarray] = []

    def fresh(self):
        return Envs(self.W, dtype=self.dtype, Ms=self.Ms, weight=self.weight)

    def lproj(self):
        return edge_env(self.dtype) if self.lpos <= 0 else self.LR[self.lpos]

    def rproj(self):
        return edge_env(self.dtype) if self.rpos >= self.N + 1 else self.LR[self.rpos]

    def makeL(self, A, k):
        ll = self.lpos
        if ll >= k:
            self.lpos = k
            return
        ll = max(ll, 0)
        L = edge_env(A[0].dtype) if ll == 0 else self.LR[ll]
        while ll < k:
            L = update_left_env(L, A[ll], self.W[ll])
            self.LR[ll + 1] = L
            ll += 1
        self.lpos = k

    def makeR(self, A, k):
        rl = self.rpos
        if rl <= k:
            self.rpos = k
            return
        rl = min(rl, self.N + 1)
        R = edge_env(A[0].dtype) if rl == self.N + 1 else self.LR[rl]
        while rl > k:
            R = update_right_env(R, A[rl - 2], self.W[rl - 2])
            self.LR[rl - 1] = R
            rl -= 1
        self.rpos = k

    def position(self, A, pos):
        """position!(P, psi, pos) = makeL!(pos-1); makeR!(pos+nsite)."""
        self.makeL(A, pos - 1)
        self.makeR(A, pos + self.nsite)
        self.m = [projector_local(A, M, pos, self.nsite) for M in self.Ms]

    def product(self, v):
        pos = self.lpos + 1
        L, R = self.lproj(), self.rproj()
        if self.nsite == 2:
            out = matvec2(L, self.W[pos - 1], self.W[pos], R, v)
        elif self.nsite == 1:
            out = matvec1(L, self.W[pos - 1], R, v)
        elif self.nsite == 0:
            out = matvec0(L, R, v)
        else:
            raise ValueError("nsite must be 0, 1 or 2")
        for m in self.m:
            out = out + self.weight * np.vdot(m, v) * m
        return out


def overlap_left(A, M, k):
    """L[b, b'] = <M|psi> over sites 1..k (psi ket, M bra; projmps2.jl:77-92)."""
    L = np.ones((1, 1), dtype=np.result_type(A[0], M[0]))
    for j in range(k):
        L = np.einsum("ab,asc,bsd->cd", L, A[j], M[j].conj())
    return L


def overlap_right(A, M, k):
    """R[a, a'] over sites k..N (projmps2.jl:108-123)."""
    N = len(A)
    R = np.ones((1, 1), dtype=np.result_type(A[0], M[0]))
    for j in range(N - 1, k - 2, -1):
        R = np.einsum("asc,bsd,cd->ab", A[j], M[j].conj(), R)
    return R


def projector_local(A, M, pos, nsite):
    """Projection m of M onto the local space at pos (1-based), such that
    ProjMPS2 v = m (m^H v) (proj_mps and product, projmps2.jl:158-224)."""
    L = overlap_left(A, M, pos - 1)
    R = overlap_right(A, M, pos + nsite)
    if nsite == 2:
        return np.einsum("ab,bsx,xtd,cd->astc", L.conj(), M[pos - 1], M[pos], R.conj())
    if nsite == 0:
        return np.einsum("ab,cb->ac", L.conj(), R.conj())
    return np.einsum("ab,bsd,cd->asc", L.conj(), M[pos - 1], R.conj())


# ---------------------------------------------------------------------------
# Krylov solvers (KrylovKit eigsolve / exponentiate restated).
# ---------------------------------------------------------------------------

@dataclass
class KrylovInfo:
    converged: int = 0
    numops: int = 0
    numiter: int = 0
    residual: float = float("nan")


def _inner(x, y):
    return np.vdot(x, y)


def _orthogonalize(w, V):
    """Two passes of classical Gram-Schmidt (full re-orthogonalisation);
    returns w' and the accumulated coefficients <v_i, w>."""
    h = np.zeros(len(V), dtype=np.result_type(w, *V) if V else w.dtype)
    for _ in range(2):
        c = np.array([_inner(v, w) for v in V])
        for ci, v in zip(c, V):
            w = w - ci * v
        h = h + c
    return w, h


def eigsolve_lanczos(apply: Callable, x0: np.ndarray, krylovdim: int = 5, maxiter: int = 2,
                     tol: float = 1e-14) -> Tuple[float, np.ndarray, KrylovInfo]:
    """Lowest (:SR) eigenpair by Lanczos with thick restart.

    Restates KrylovKit ``eigsolve(A, x0, 1, :SR; ishermitian=true, krylovdim,
    maxiter, tol, eager=false)`` as called at src/base/solver.jl:36:
      * factorisation grows one vector per operator application until it has
        ``krylovdim`` vectors (or the residual norm drops to ``tol``);
      * Ritz pairs of the projected matrix, residual |beta * U[K-1, i]|;
      * stop when the lowest Ritz pair is converged or ``numiter == maxiter``;
      * otherwise keep ``(3*krylovdim + 2*converged) // 5`` Ritz vectors plus
        the residual direction (arrowhead projected matrix) and expand again.
    Returns the Ritz value and the (unit) Ritz vector (solver.jl:42)."""
    shape = x0.shape
    x = x0.reshape(-1)
    beta0 = np.linalg.norm(x)
    V = [x / beta0]
    Hm = np.zeros((krylovdim + 1, krylovdim + 1), dtype=np.float64)
    info = KrylovInfo(numiter=1)
    K = 1
    # initialise: first operator application
    w = apply(V[0].reshape(shape)).reshape(-1)
    info.numops += 1
    w, h = _orthogonalize(w, V)
    Hm[0, 0] = h[0].real
    r = w
    beta = np.linalg.norm(r)
    while True:
        if K == krylovdim or beta <= tol:
            H = Hm[:K, :K]
            H = 0.5 * (H + H.T)
            theta, U = np.linalg.eigh(H)
            res = np.abs(beta * U[K - 1, :])
            conv = 0
            while conv < K and res[conv] <= tol:
                conv += 1
            info.converged = conv
            info.residual = res[0]
            if conv >= 1 or beta <= tol or info.numiter >= maxiter:
                y = sum(U[i, 0] * V[i] for i in range(K))
                y = y / np.linalg.norm(y)
                return float(theta[0]), y.reshape(shape), info
            # thick restart
            keep = (3 * krylovdim + 2 * conv) // 5
            keep = max(1, min(keep, K - 1))
            newV = [sum(U[i, j] * V[i] for i in range(K)) for j in range(keep)]
            Hn = np.zeros_like(Hm)
            for j in range(keep):
                Hn[j, j] = theta[j]
                Hn[keep, j] = Hn[j, keep] = beta * U[K - 1, j]
            V = newV + [r / beta]
            Hm = Hn
            K = keep + 1
            info.numiter += 1
        else:
            V.append(r / beta)
            Hm[K, K - 1] = Hm[K - 1, K] = beta
            K += 1
        # expand: apply the operator to the newest basis vector; the projected
        # matrix keeps the Lanczos recurrence values (alpha on the diagonal,
        # beta = |r| off it); the other Gram-Schmidt coefficients only
        # re-orthogonalise.
        w = apply(V[-1].reshape(shape)).reshape(-1)
        info.numops += 1
        w, h = _orthogonalize(w, V)
        Hm[K - 1, K - 1] = h[K - 1].real
        r = w
        beta = np.linalg.norm(r)


def _small_expm_e1(H: np.ndarray, t: complex) -> np.ndarray:
    """exp(t H) e_1 for a small real symmetric H (eigen-decomposition)."""
    theta, U = np.linalg.eigh(0.5 * (H + H.T))
    return U @ (np.exp(t * theta) * U[0, :])


def exponentiate_lanczos(apply: Callable, t: complex, x0: np.ndarray, krylovdim: int = 30,
                         maxiter: int = 100, tol: float = 1e-12) -> Tuple[np.ndarray, KrylovInfo]:
    """exp(t A) x0 for Hermitian A by a Lanczos exponential integrator.

    Restates KrylovKit ``exponentiate(A, t, x0; ishermitian=true, tol,
    krylovdim, maxiter, eager=true)`` (src/base/solver.jl:80): grow the
    Lanczos basis of the current vector; after every expansion (eager) take
    y = exp(tau H_K) e_1 and the a-posteriori error |beta * y_K| * |x|;
    accept the remaining step when that error is below tol*|tau|/|t|;
    at ``krylovdim`` without acceptance, take the largest step tau' that
    meets the error bound (found by halving) and restart from the
    propagated vector.  The step-size controller details are KrylovKit's
    own and are parity-unpinned; converged results equal exp(tA)x0 to tol."""
    shape = x0.shape
    x = x0.reshape(-1).astype(np.complex128 if np.iscomplexobj(t) or np.iscomplexobj(x0) else np.float64)
    info = KrylovInfo()
    total = t
    remaining = 1.0  # fraction of t still to go
    last = 1.0       # last accepted sub-step (fraction of t)
    while remaining > 0 and info.numiter < maxiter:
        info.numiter += 1
        nx = np.linalg.norm(x)
        if nx == 0:
            break
        V = [x / nx]
        Hm = np.zeros((krylovdim + 1, krylovdim + 1))
        K = 0
        while True:
            w = apply(V[K].reshape(shape)).reshape(-1)
            info.numops += 1
            w, h = _orthogonalize(w, V)
            Hm[K, K] = h[K].real
            K += 1
            beta = np.linalg.norm(w)
            tau = remaining * total
            y = _small_expm_e1(Hm[:K, :K], tau)
            err = abs(beta * y[K - 1]) * nx
            if err <= tol * remaining or beta <= 1e-14 * max(1.0, np.abs(Hm[:K, :K]).max()):
                x = nx * sum(y[i] * V[i] for i in range(K))
                remaining = 0.0
                break
            if K == krylovdim:
                # largest sub-step (halving from twice the last accepted one)
                # whose error estimate meets the per-unit-time tolerance
                frac = min(remaining, 2.0 * last)
                while True:
                    yy = _small_expm_e1(Hm[:K, :K], frac * total)
                    if abs(beta * yy[K - 1]) * nx <= tol * frac or frac <= 1e-12:
                        break
                    frac *= 0.5
                x = nx * sum(yy[i] * V[i] for i in range(K))
                remaining -= frac
                last = frac
                break
            V.append(w / beta)
            Hm[K, K - 1] = Hm[K - 1, K] = beta
    info.converged = 1 if remaining == 0 else 0
    return x.reshape(shape), info


# ---------------------------------------------------------------------------
# SVD truncation (NDTensors truncate!!) and replacebond!.
# ---------------------------------------------------------------------------

@dataclass
class TruncResult:
    n: int
    truncerr: float
    docut: float


def truncate_weights(P: np.ndarray, maxdim: int, mindim: int, cutoff: float) -> TruncResult:
    """NDTensors ``truncate!!`` with the relative cutoff (ITensors' default):
      * zero negative trailing weights;
      * drop from the end while n > maxdim (their weight is the error);
      * then drop while truncerr + P[n] <= cutoff * sum(P) and n > mindim;
      * truncerr /= sum(P);
      * docut = (P[n] + P[n+1]) / 2, bumped by 1e-3 P[n] when the two are
        within 1e-3 P[n] (used by block-sparse tensors to drop whole
        near-degenerate multiplets).
    P must be sorted in descending order."""
    P = np.array(P, dtype=np.float64)
    origm = len(P)
    for k in range(origm - 1, -1, -1):
        if P[k] >= 0:
            break
        P[k] = 0.0
    if origm == 1:
        return TruncResult(1, 0.0, abs(P[0]) / 2)
    n = origm
    truncerr = 0.0
    maxdim = min(maxdim, origm)
    mindim = max(1, min(mindim, maxdim))
    while n > maxdim:
        truncerr += P[n - 1]
        n -= 1
    scale = P.sum()
    if scale == 0.0:
        scale = 1.0
    while n > mindim and truncerr + P[n - 1] <= cutoff * scale:
        truncerr += P[n - 1]
        n -= 1
    truncerr /= scale
    n = max(n, 1)
    docut = 0.0
    if n < origm:
        docut = (P[n - 1] + P[n]) / 2
        if abs(P[n - 1] - P[n]) < 1e-3 * P[n - 1]:
            docut += 1e-3 * P[n - 1]
    return TruncResult(n, truncerr, docut)


@dataclass
class Spectrum:
    eigs: np.ndarray
    truncerr: float


def svd_truncate(M: np.ndarray, maxdim: int, mindim: int, cutoff: float):
    """Dense truncated SVD (LAPACK gesdd, ITensors svd_alg divide_and_conquer)
    keeping the first n singular triplets chosen by truncate_weights."""
    U, S, Vh = np.linalg.svd(M, full_matrices=False)
    P = S * S
    tr = truncate_weights(P, maxdim, mindim, cutoff)
    n = tr.n
    return U[:, :n], S[:n], Vh[:n, :], Spectrum(P[:n].copy(), tr.truncerr)


def noiseterm_factor(env: "Envs", phi: np.ndarray, ortho: str) -> np.ndarray:
    """Factor X of the density-matrix perturbation of the two-site update,
    noiseterm(PH, phi, ortho) = X X^H over the indices that stay on the
    orthonormal side (ITensorMPS 0.3 noiseterm(::AbstractProjMPO), called at
    src/mps/update_site.jl:60-62; the in-tree mirror for coupling models is
    src/mps/projcouplingmodel.jl:391-492):
      left:  X[(a' s'), (w t b)] = sum L[a, w0, a'] W1[w0, s', s, w] phi[a, s, t, b]
      right: X[(a s w), (t' b')] = sum phi[a, s, t, b] W2[w, t', t, w2] R[b, w2, b']
    Rows (left) or columns (right) are the indices of the kept factor; the
    perturbation in the convention of replacebond below is X X^H (left) and
    X^T conj(X) (right)."""
    if env.nsite != 2:
        raise ValueError("noise term only defined for 2-site ProjMPO")
    pos = env.lpos + 1
    a, s, t, c = phi.shape
    if ortho == "left":
        T = np.tensordot(env.lproj(), phi, axes=([0], [0]))        # [w0, a', s, t, b]
        T = np.tensordot(T, env.W[pos - 1], axes=([0, 2], [0, 2]))  # [a', t, b, s', w]
        return T.transpose(0, 3, 4, 1, 2).reshape(a * s, -1)
    if ortho == "right":
        T = np.tensordot(phi, env.W[pos], axes=([2], [2]))          # [a, s, b, w, t', w2]
        T = np.tensordot(T, env.rproj(), axes=([2, 5], [0, 1]))     # [a, s, w, t', b']
        return T.reshape(-1, t * c)
    raise ValueError(f"In noiseterm, got ortho = {ortho}, only supports `left` and `right`")


def noiseterm(env: "Envs", phi: np.ndarray, ortho: str) -> np.ndarray:
    if hasattr(env, "noiseterm"):  # ProjCouplingModel / ProjMPOSum2 (tn_oracle_cm)
        return env.noiseterm(phi, ortho)
    X = noiseterm_factor(env, phi, ortho)
    return X @ X.conj().T if ortho == "left" else X.T @ X.conj()


def replacebond_perturbed(A: List[np.ndarray], b: int, phi: np.ndarray, maxdim: int, mindim: int,
                          cutoff: float, ortho: str, normalize: bool, drho: np.ndarray) -> Spectrum:
    """replacebond!(psi, b, phi; eigen_perturbation=drho, ortho): the
    density-matrix factorisation (ITensors 0.9 factorize_eigen).  The
    orthonormal factor is the top-n eigenvectors of rho = phi phi^H + drho
    (left; rho = phi^T conj(phi) + drho on the right), the other factor is its
    exact projection of phi, normalised; the spectrum is that of rho."""
    a, s, t, c = phi.shape
    M = phi.reshape(a * s, t * c)
    rho = M @ M.conj().T if ortho == "left" else M.T @ M.conj()
    rho = rho + drho
    rho = 0.5 * (rho + rho.conj().T)
    w, V = np.linalg.eigh(rho)
    w, V = w[::-1], V[:, ::-1]
    P = np.maximum(w, 0.0)
    tr = truncate_weights(P, maxdim, mindim, cutoff)
    n = tr.n
    U = V[:, :n]
    if ortho == "left":
        R = U.conj().T @ M
        if normalize:
            R = R / np.linalg.norm(R)
        A[b] = U.reshape(a, s, n)
        A[b + 1] = R.reshape(n, t, c)
    elif ortho == "right":
        Lf = M @ U.conj()
        if normalize:
            Lf = Lf / np.linalg.norm(Lf)
        A[b] = Lf.reshape(a, s, n)
        A[b + 1] = U.T.reshape(n, t, c)
    else:
        raise ValueError("ortho must be 'left' or 'right'")
    return Spectrum(P[:n].copy(), tr.truncerr)


def replacebond(A: List[np.ndarray], b: int, phi: np.ndarray, maxdim: int, mindim: int,
                cutoff: float, ortho: str, normalize: bool = True) -> Spectrum:
    """replacebond!(psi, b, phi; ortho) for 0-based bond b (sites b, b+1):
    ortho="left" -> A_b = U, A_{b+1} = S V (normalised);
    ortho="right" -> A_{b+1} = V, A_b = U S (normalised)."""
    a, s, t, c = phi.shape
    U, S, Vh, spec = svd_truncate(phi.reshape(a * s, t * c), maxdim, mindim, cutoff)
    n = S.size
    if normalize:
        S = S / np.linalg.norm(S)
    if ortho == "left":
        A[b] = U.reshape(a, s, n)
        A[b + 1] = (S[:, None] * Vh).reshape(n, t, c)
    elif ortho == "right":
        A[b] = (U * S[None, :]).reshape(a, s, n)
        A[b + 1] = Vh.reshape(n, t, c)
    else:
        raise ValueError("ortho must be 'left' or 'right'")
    return spec


def entropy(p: np.ndarray) -> float:
    """_entropy (src/mps/measure.jl:4-20)."""
    p = np.asarray(p, dtype=np.float64)
    s = p.sum()
    if abs(s - 1.0) > 100 * FLOAT64_THRESHOLD:
        p = p / s
    return float(-sum(x * math.log(x) for x in p if x > 0.0))


# ---------------------------------------------------------------------------
# Sweeps: _update_two_site!, fullsweep!, dmrg!, tdvpsweep!.
# ---------------------------------------------------------------------------

@dataclass
class SweepData:
    """src/mps/sweep.jl:26-35."""
    sweepcount: int = 0
    maxchi: List[int] = field(default_factory=list)
    energy: List[float] = field(default_factory=list)
    entropy: List[float] = field(default_factory=list)
    maxtruncerr: List[float] = field(default_factory=list)
    lasteigs: List[np.ndarray] = field(default_factory=list)


@dataclass
class DMRGParams:
    """src/mps/dmrg.jl:26-102 (scalars broadcast; lengths must agree)."""
    maxdim: List[int]
    nsweeps: List[int]
    cutoff: List[float]
    noise: List[float]
    noisedecay: List[float]
    disable_noise_after: List[int]

    @staticmethod
    def make(maxdim, nsweeps, cutoff=FLOAT64_THRESHOLD, noise=0.0, noisedecay=1.0,
             disable_noise_after=-1):
        n = len(nsweeps)
        f = lambda x, cast: list(x) if isinstance(x, (list, tuple, np.ndarray)) else [cast(x)] * n
        p = DMRGParams(list(maxdim), list(nsweeps), f(cutoff, float), f(noise, float),
                       f(noisedecay, float), f(disable_noise_after, int))
        if not (len(p.nsweeps) == len(p.maxdim) == len(p.cutoff) == len(p.noise)
                == len(p.noisedecay) == len(p.disable_noise_after)):
            raise ValueError("`DMRGParams()` :: Size mismatch in input vectors !!")
        return p


class StateEnvs:
    """{psi::MPS, PH} container (src/mps/state_envs.jl:18-60)."""

    def __init__(self, psi: List[np.ndarray], Ws: List[np.ndarray], Ms=None, weight: Optional[float] = None):
        """StateEnvs(psi, H) or StateEnvs(psi, H, Ms; weight) (state_envs.jl:54-106)."""
        if Ms is not None and (weight is None or weight <= 0.0):
            raise ValueError(f"`weight` parameter should be > 0.0 (value passed was `weight={weight}`)")
        self.psi = [a.copy() for a in psi]
        from oracle import tn_oracle_cm as CM
        if isinstance(Ws, CM.CMModel):      # ProjCouplingModel(_MPS), state_envs.jl:136-175
            self.PH = CM.ProjCM(Ws, Ms=Ms, weight=weight or 0.0)
        elif len(Ws) and not isinstance(Ws[0], np.ndarray):  # ProjMPOSum2, state_envs.jl:108-134
            self.PH = CM.ProjSum(Ws, dtype=psi[0].dtype, Ms=Ms, weight=weight or 0.0)
        else:
            self.PH = Envs(Ws, dtype=psi[0].dtype, Ms=Ms, weight=weight or 0.0)
        self.ortho_center: Optional[int] = None  # 1-based when known

    def __len__(self):
        return len(self.psi)


def _ensure_ortho_first(env: StateEnvs):
    if env.ortho_center != 1:
        env.psi = right_orthogonalize(env.psi)
        env.ortho_center = 1
        env.PH = env.PH.fresh()


def update_two_site(env: StateEnvs, pos: int, ortho: str, *, solver: str = "eig",
                    time_step=None, maxdim=2 ** 62, mindim=1, cutoff=FLOAT64_THRESHOLD,
                    normalize=True, krylovdim=5, maxiter=2, tol=1e-14,
                    reverse_step=False, noise=0.0):
    """_update_two_site! (src/mps/update_site.jl:27-90); pos is 1-based."""
    N = len(env)
    if not (0 < pos < N):
        raise ValueError("position out of range")
    A = env.psi
    env.PH.nsite = 2
    phi = np.tensordot(A[pos - 1], A[pos], axes=([2], [0]))
    env.PH.position(A, pos)
    if solver == "eig":
        energy, phi, _ = eigsolve_lanczos(env.PH.product, phi, krylovdim, maxiter, tol)
    else:
        phi, _ = exponentiate_lanczos(env.PH.product, time_step, phi, krylovdim, maxiter, tol)
        energy = float("nan")
    if normalize:
        phi = phi / np.linalg.norm(phi)
    if math.isnan(energy):
        energy = float(np.real(np.vdot(phi, env.PH.product(phi))))
    if abs(noise) > FLOAT64_THRESHOLD:
        drho = noise * noiseterm(env.PH, phi, ortho)
        spec = replacebond_perturbed(A, pos - 1, phi, maxdim, mindim, cutoff, ortho, normalize, drho)
    else:
        spec = replacebond(A, pos - 1, phi, maxdim, mindim, cutoff, ortho, normalize)
    env.ortho_center = pos + 1 if ortho == "left" else pos
    halfsweep_done = (pos == 1 and ortho == "right") or (pos == N - 1 and ortho == "left")
    if reverse_step and not halfsweep_done:
        pos1 = pos + 1 if ortho == "left" else pos
        phi0 = A[pos1 - 1]
        env.PH.nsite = 1
        env.PH.position(A, pos1)
        phi0, _ = exponentiate_lanczos(env.PH.product, -time_step, phi0, krylovdim, maxiter, tol)
        if normalize:
            phi0 = phi0 / np.linalg.norm(phi0)
        energy = float(np.real(np.vdot(phi0, env.PH.product(phi0))))
        A[pos1 - 1] = phi0
    return energy, spec.truncerr, spec.eigs


def fullsweep(env: StateEnvs, swdata: SweepData, **kw):
    """fullsweep!(sysenv, solver, 2, swdata; kwargs...) (src/mps/sweep.jl:94-196)."""
    _ensure_ortho_first(env)
    N = len(env)
    energy = float("nan")
    maxtruncerr = 0.0
    swdata.sweepcount += 1
    lasteigs: List[np.ndarray] = [None] * (N - 1)
    for bond in range(1, N):
        energy, err, _ = update_two_site(env, bond, "left", **kw)
        maxtruncerr = max(err, maxtruncerr)
    for bond in range(N - 1, 0, -1):
        energy, err, eigs = update_two_site(env, bond, "right", **kw)
        lasteigs[bond - 1] = eigs
        maxtruncerr = max(err, maxtruncerr)
    swdata.maxchi.append(max(a.shape[2] for a in env.psi))
    swdata.energy.append(energy)
    mid = lasteigs[N // 2 - 1]
    swdata.entropy.append(entropy(mid / mid.sum()))
    swdata.maxtruncerr.append(maxtruncerr)
    swdata.lasteigs = lasteigs
    if swdata.sweepcount > 1:
        return swdata.energy[-1] - swdata.energy[-2], swdata.entropy[-1] - swdata.entropy[-2]
    return float("nan"), float("nan")


def dmrg(env: StateEnvs, params: DMRGParams, *, krylovdim=5, maxiter=2, tol=1e-14,
         energyErrGoal=None, normalize=True) -> SweepData:
    """dmrg!(sysenv, params, 2) (src/mps/dmrg.jl:148-227), with the noise
    schedule of dmrg.jl:196-205."""
    sw = SweepData()
    for ii in range(len(params.nsweeps)):
        noise = params.noise[ii]
        for jj in range(1, params.nsweeps[ii] + 1):
            enerr, _ = fullsweep(env, sw, maxdim=params.maxdim[ii], cutoff=params.cutoff[ii],
                                 krylovdim=krylovdim, maxiter=maxiter, tol=tol,
                                 normalize=normalize, noise=noise)
            if energyErrGoal is not None and abs(enerr) < abs(energyErrGoal) and abs(noise) < FLOAT64_THRESHOLD:
                break
            if jj == params.disable_noise_after[ii]:
                noise = 0.0
            noise /= params.noisedecay[ii]
            if noise < 100 * FLOAT64_THRESHOLD:
                noise = 0.0
    return sw


def tdvpsweep(env: StateEnvs, swdata: SweepData, time_step: complex, *, maxdim=2 ** 62,
              cutoff=FLOAT64_THRESHOLD, krylovdim=30, maxiter=100, tol=1e-12):
    """tdvpsweep!(engine, time_step; nsite=2) (src/mps/tdvp.jl:247-277):
    fullsweep! with exp_solver, half time steps and the one-site reverse step."""
    return fullsweep(env, swdata, solver="exp", time_step=0.5 * time_step, maxdim=maxdim,
                     cutoff=cutoff, krylovdim=krylovdim, maxiter=maxiter, tol=tol,
                     reverse_step=True)


# ---------------------------------------------------------------------------
# Small exact references.
# ---------------------------------------------------------------------------

def mps_energy(A: List[np.ndarray], Ws: List[np.ndarray]) -> float:
    """<psi|H|psi>/<psi|psi> by a full left-environment contraction."""
    L = edge_env(A[0].dtype)
    for Ak, Wk in zip(A, Ws):
        L = update_left_env(L, Ak, Wk)
    nrm = 1.0
    Ln = np.ones((1, 1))
    for Ak in A:
        Ln = np.einsum("ab,asc,bsd->cd", Ln, Ak, Ak.conj())
    nrm = Ln.reshape(-1)[0]
    return float(np.real(L.reshape(-1)[0] / nrm))
