"""FP64 CPU oracle for sums of term strings on an MPS: CouplingModel,
ProjCouplingModel and ProjMPOSum2 (numpy).

TEST INFRASTRUCTURE ONLY.  This module is a checker for the HIP path; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
may import it.  The product path never routes through it.

What it restates (reference = TenNetLib.jl 0.2.2):

* ``CouplingModel(os, sites)`` (``src/base/couplingmodel.jl:153-241``) with
  the OpStrings clean-up it calls (``removeIdsZeros``, ``bosonize``,
  ``mergeterms``: ``src/base/opstrings.jl:161-237``,
  ``src/base/fermions.jl:30-96``), the per-term factors
  (``couplingmodel.jl:175-185``, ``_add_oplinks!``,
  ``src/base/helper_internal_funcs.jl:41-63``) and the merge of strings on one
  support (``_chunksum_cm_terms``, ``couplingmodel.jl:65-131``);
  ``CouplingModel(os, mpo)`` and ``CouplingModel(mpos...)``
  (``couplingmodel.jl:292-356``).
* ``ProjCouplingModel`` (``src/mps/projcouplingmodel.jl``): environments as
  dictionaries of per-id tensors (``_makeL!``/``_makeR!``, ``:124-303``, with
  completed ids summed into one tensor), ``position!`` (``:311-315``), the
  per-id product (``_contract``/``product``, ``:319-375``) and the per-id
  noise term (``noiseterm``, ``:391-492``).
* ``ProjMPOSum2`` (``src/mps/projmposum2.jl:80-140``): the sum of
  ``ProjMPO`` products and noise terms.

Tensors: factors F[l, s', s, r] with unit links at the ends of a string,
environments L[a, w, a'] / R[b, w, b'] (ket, link, bra) as in tn_oracle, and a
completed-terms environment as a matrix [ket, bra].

Pins (tests/test_oracle_cm.py): products equal the projected dense
Hamiltonian P^H H P; the merged factors of a Heisenberg bond have a closed
form; the left/right noise terms of one-string models have closed forms;
the noise of CouplingModel(mpo) equals the ProjMPO noise term.
"""

from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from oracle import tn_oracle as O

FLOAT64_THRESHOLD = O.FLOAT64_THRESHOLD

_OPS = {
    "Id": np.eye(2), "Sz": np.diag([0.5, -0.5]),
    "S+": np.array([[0.0, 1.0], [0.0, 0.0]]), "S-": np.array([[0.0, 0.0], [1.0, 0.0]]),
    "Sx": np.array([[0.0, 0.5], [0.5, 0.0]]),
}


def op(name: str) -> np.ndarray:
    """op("A * B", s) = A B (several operators on one site, fermions.jl:79-91)."""
    out = np.eye(2)
    for part in name.split(" * "):
        out = out @ _OPS[part.strip()]
    return out


# ---------------------------------------------------------------------------
# CouplingModel
# ---------------------------------------------------------------------------

def clean_terms(terms, N):
    """removeIdsZeros -> bosonize -> mergeterms for spin-1/2 strings.
    terms: iterable of (coef, [(name, site), ...]) -> list of
    (coef, ((name, site), ...)) with one operator per site, sites sorted."""
    acc: Dict[tuple, float] = {}
    order: List[tuple] = []
    for coef, ops in terms:
        ops = [x for x in ops if x[0] != "Id"]                     # removeIds (opstrings.jl:86-96)
        if not abs(coef) > FLOAT64_THRESHOLD:                     # removeIdsZeros (:161-165)
            continue
        sites: Dict[int, str] = {}
        for name, i in ops:                                        # bosonize, no fermions (:79-91)
            sites[i] = name if i not in sites else sites[i] + " * " + name
        key = tuple((sites[i], i) for i in sorted(sites))
        for _, i in key:
            if i < 1 or i > N:
                raise ValueError("`CouplingModel()`: Out-of-bounds error in OpStrings !!")
        if key not in acc:
            order.append(key)
            acc[key] = 0.0
        acc[key] += coef                                           # mergeterms (:213-226)
    out = [(acc[k], k) for k in order if abs(acc[k]) > FLOAT64_THRESHOLD]
    out.sort(key=lambda t: (t[1][0][1], len(t[1])))               # isless (:68-77)
    return out


def string_factors(coef, key):
    n = len(key)
    fs = [abs(coef) ** (1.0 / n) * op(name) for name, _ in key]
    fs[0] = np.sign(coef) * fs[0]
    return [f.reshape(1, 2, 2, 1) for f in fs]


def _dsum(strings):
    """_directsum over the links of strings on one support (helper_internal_funcs.jl:73-110)."""
    n = len(strings[0])
    res = []
    for j in range(n):
        parts = [s[j] for s in strings]
        if n == 1:
            res.append(sum(parts))
            continue
        if j == 0:
            res.append(np.concatenate(parts, axis=3))
        elif j == n - 1:
            res.append(np.concatenate(parts, axis=0))
        else:
            L = sum(p.shape[0] for p in parts)
            R = sum(p.shape[3] for p in parts)
            T = np.zeros((L, 2, 2, R))
            a = b = 0
            for p in parts:
                T[a:a + p.shape[0], :, :, b:b + p.shape[3]] = p
                a += p.shape[0]
                b += p.shape[3]
            res.append(T)
    return res


def _sum_and_compress(strings):
    """couplingmodel.jl:80-104: direct sum, then SVD sweeps left and right with
    cutoff 1e-15; S normalised, absorbed in the sweep direction, sqrt(|S|)
    on both sides."""
    T = _dsum(strings)
    n = len(T)
    for direction in ("left", "right"):
        for b in (range(n - 1) if direction == "left" else reversed(range(n - 1))):
            l = T[b].shape[0]
            r = T[b + 1].shape[3]
            M = np.einsum("astu,uxyb->astxyb", T[b], T[b + 1]).reshape(l * 4, 4 * r)
            U, S, Vh = np.linalg.svd(M, full_matrices=False)
            k = O.truncate_weights(S ** 2, 2 ** 62, 1, FLOAT64_THRESHOLD).n
            U, S, Vh = U[:, :k], S[:k], Vh[:k, :]
            nrm = np.sqrt(np.sum(S ** 2))
            S = S / nrm
            if direction == "left":
                Vh = np.diag(S) @ Vh
            else:
                U = U @ np.diag(S)
            T[b] = np.sqrt(nrm) * U.reshape(l, 2, 2, k)
            T[b + 1] = np.sqrt(nrm) * Vh.reshape(k, 2, 2, r)
    return T


def _chunksum(strings, chunksize=12):
    if len(strings) == 1:
        return strings[0]
    if len(strings) <= chunksize:
        return _sum_and_compress(strings)
    chunks = [strings[i:i + chunksize] for i in range(0, len(strings), chunksize)]
    return _chunksum([_chunksum(c, chunksize) for c in chunks], chunksize)


class CMModel:
    """CouplingModel: ``ids`` is a list of (support, factors)."""

    def __init__(self, N: int, ids: List[Tuple[Tuple[int, ...], List[np.ndarray]]]):
        self.N = N
        self.ids = ids

    @classmethod
    def from_terms(cls, terms, N, merge=True, mpo=None):
        """CouplingModel(os, sites; merge) (couplingmodel.jl:153-241), plus
        the MPO as one more id for CouplingModel(os, mpo) (:292-322)."""
        groups: Dict[tuple, list] = {}
        gorder = []
        for coef, key in clean_terms(terms, N):
            pos = tuple(i for _, i in key)
            fs = string_factors(coef, key)
            if pos not in groups:
                groups[pos] = [fs]
                gorder.append(pos)
            elif len(pos) == 1:
                groups[pos][0] = [groups[pos][0][0] + fs[0]]      # .+= (couplingmodel.jl:189-191)
            else:
                groups[pos].append(fs)
        ids = []
        for pos in gorder:
            if merge:
                ids.append((pos, _chunksum(groups[pos])))
            else:
                ids += [(pos, fs) for fs in groups[pos]]
        if mpo is not None:
            ids.append((tuple(range(1, N + 1)), [np.asarray(W, float) for W in mpo]))
        return cls(N, ids)

    @classmethod
    def from_mpos(cls, *mpos):
        """CouplingModel(mpos...) (couplingmodel.jl:331-356)."""
        N = len(mpos[0])
        return cls(N, [(tuple(range(1, N + 1)), [np.asarray(W, float) for W in m]) for m in mpos])

    def dense(self, d=2) -> np.ndarray:
        """Sum of the strings as a dense d^N x d^N matrix (small N)."""
        H = 0.0
        for pos, fs in self.ids:
            mats = []
            for i in range(1, self.N + 1):
                mats.append(fs[pos.index(i)] if i in pos else None)
            T = np.ones((1, 1, 1))  # [S', S, link]
            for i, F in enumerate(mats):
                if F is None:
                    T = np.einsum("ABw,xy->AxByw", T, np.eye(d))
                    a, x, b, y, w = T.shape
                    T = T.reshape(a * x, b * y, w)
                else:
                    T = np.einsum("ABw,wxyz->AxByz", T, F)
                    a, x, b, y, z = T.shape
                    T = T.reshape(a * x, b * y, z)
            H = H + T[:, :, 0]
        return H


# ---------------------------------------------------------------------------
# ProjCouplingModel
# ---------------------------------------------------------------------------

_DONE = "done"


class ProjCM:
    """ProjCouplingModel with lpos/rpos caching (projcouplingmodel.jl:14-315),
    plus the penalty projectors of ProjCouplingModel_MPS."""

    def __init__(self, model: CMModel, Ms=None, weight: float = 0.0):
        self.model = model
        self.N = model.N
        self.M: List[Dict[int, np.ndarray]] = [dict() for _ in range(self.N + 2)]
        self.ends: Dict[int, Tuple[int, int]] = {}
        for q, (pos, fs) in enumerate(model.ids):
            for i, F in zip(pos, fs):
                self.M[i][q] = F
            self.ends[q] = (pos[0], pos[-1])
        self.LR: List[Optional[dict]] = [None] * (self.N + 2)
        self.lpos = 0
        self.rpos = self.N + 1
        self.nsite = 2
        self.Ms = Ms or []
        self.weight = weight
        self.m: List[np.ndarray] = []

    def fresh(self):
        return ProjCM(self.model, self.Ms, self.weight)

    def lproj(self) -> dict:
        return {} if self.lpos <= 0 else self.LR[self.lpos]

    def rproj(self) -> dict:
        return {} if self.rpos >= self.N + 1 else self.LR[self.rpos]

    def _step_left(self, L: dict, A: np.ndarray, site: int) -> dict:
        """_makeL! body (projcouplingmodel.jl:124-207) for one site."""
        new, local = {}, None
        Ac = A.conj()
        for q in set(L) | set(self.M[site]):
            if q in L and q in self.M[site]:
                T = np.einsum("awx,asb,wtsv,xtc->bvc", L[q], A, self.M[site][q], Ac)
            elif q in L:
                if q == _DONE:
                    T = np.einsum("ax,asb,xsc->bc", L[q], A, Ac)
                else:
                    T = np.einsum("awx,asb,xsc->bwc", L[q], A, Ac)
            else:  # starts here: the left link contracts bra with ket
                T = np.einsum("asb,tsv,atc->bvc", A, self.M[site][q][0], Ac)
            if q == _DONE or self.ends[q][1] == site:  # order 2: completed
                T = T.reshape(T.shape[0], T.shape[-1]) if T.ndim == 3 else T
                local = T if local is None else local + T
            else:
                new[q] = T
        if local is not None:
            new[_DONE] = local
        return new

    def _step_right(self, R: dict, A: np.ndarray, site: int) -> dict:
        """_makeR! body (projcouplingmodel.jl:220-299) for one site."""
        new, local = {}, None
        Ac = A.conj()
        for q in set(R) | set(self.M[site]):
            if q in R and q in self.M[site]:
                T = np.einsum("asb,wtsv,bvc,xtc->awx", A, self.M[site][q], R[q], Ac)
            elif q in R:
                if q == _DONE:
                    T = np.einsum("asb,bc,xsc->ax", A, R[q], Ac)
                else:
                    T = np.einsum("asb,bvc,xsc->avx", A, R[q], Ac)
            else:  # ends here: the right link contracts bra with ket
                T = np.einsum("asb,wts,xtb->awx", A, self.M[site][q][..., 0], Ac)
            if q == _DONE or self.ends[q][0] == site:
                T = T.reshape(T.shape[0], T.shape[-1]) if T.ndim == 3 else T
                local = T if local is None else local + T
            else:
                new[q] = T
        if local is not None:
            new[_DONE] = local
        return new

    def makeL(self, A, k):
        ll = self.lpos
        if ll >= k:
            self.lpos = k
            return
        ll = max(ll, 0)
        L = self.lproj() if ll > 0 else {}
        while ll < k:
            L = self._step_left(L, A[ll], ll + 1)
            self.LR[ll + 1] = L
            ll += 1
        self.lpos = k

    def makeR(self, A, k):
        rl = self.rpos
        if rl <= k:
            self.rpos = k
            return
        rl = min(rl, self.N + 1)
        R = self.LR[rl] if rl <= self.N else {}
        while rl > k:
            R = self._step_right(R, A[rl - 2], rl - 1)
            self.LR[rl - 1] = R
            rl -= 1
        self.rpos = k

    def position(self, A, pos):
        self.makeL(A, pos - 1)
        self.makeR(A, pos + self.nsite)
        self.m = [O.projector_local(A, M, pos, self.nsite) for M in self.Ms]

    # -- per-id pieces -----------------------------------------------------

    def _sites(self):
        p = self.lpos + 1
        return list(range(p, p + self.nsite))

    def _factor(self, q, site, wl):
        """Factor of id q at site, or its identity pass-through of link dim wl."""
        if q in self.M[site]:
            return self.M[site][q]
        return np.einsum("wu,ts->wtsu", np.eye(wl), np.eye(2))

    @staticmethod
    def _apply_left(q, L, v):
        """T[a', w, s..., b]: left piece of open id q applied to v[a, s..., b]
        (the identity with a unit link when q has not acted left of the sites)."""
        if q in L:
            return np.tensordot(L[q], v, axes=([0], [0])).transpose(1, 0, *range(2, v.ndim + 1))
        return v[:, None]

    def product(self, v):
        """sum_id contract(v, tensors of id) (projcouplingmodel.jl:319-375)."""
        L, R = self.lproj(), self.rproj()
        sites = self._sites()
        ns = len(sites)
        out = np.zeros_like(v, dtype=np.result_type(v, 1.0))
        ids = set(L) | set(R)
        for s in sites:
            ids |= set(self.M[s])
        for q in ids:
            if q == _DONE:
                continue
            T = self._apply_left(q, L, v)          # [a', w, s1..sn, b]
            for j, s in enumerate(sites):
                F = self._factor(q, s, T.shape[1])  # [w, s', s, u]
                T = np.tensordot(T, F, axes=([1, 2 + j], [0, 2]))  # [a', s.., b, s', u]
                T = np.moveaxis(T, -1, 1)
                T = np.moveaxis(T, -1, 2 + j)
            # T[a', u, s'..., b]
            if q in R:
                T = np.tensordot(T, R[q], axes=([1, T.ndim - 1], [1, 0]))  # [a', s'..., b']
            else:
                T = T[:, 0]
            out = out + T
        if _DONE in L:
            out = out + np.tensordot(L[_DONE], v, axes=([0], [0]))
        if _DONE in R:
            out = out + np.tensordot(v, R[_DONE], axes=([ns + 1], [0]))
        for m in self.m:
            out = out + self.weight * np.vdot(m, v) * m
        return out

    def noiseterm_factors(self, phi, ortho):
        """Per-id factors X_id of noiseterm (projcouplingmodel.jl:391-492):
        left X[(a', s'), (u, t, b)], right X[(a, s, w), (t', b')]."""
        if self.nsite != 2:
            raise ValueError("noise term only defined for 2-site ProjMPO")
        p = self.lpos + 1
        a, s, t, c = phi.shape
        Xs = []
        if ortho == "left":
            L = self.lproj()
            for q in set(L) | set(self.M[p]):
                if q == _DONE:
                    X = np.tensordot(L[q], phi, axes=([0], [0]))             # [a', s, t, b]
                    Xs.append(X.reshape(a * s, t * c))
                    continue
                T = self._apply_left(q, L, phi)                             # [a', w, s, t, b]
                F = self._factor(q, p, T.shape[1])
                T = np.einsum("xwstb,wrsu->xrutb", T, F)                     # [a', s', u, t, b]
                Xs.append(T.reshape(a * s, -1))
            return Xs
        if ortho == "right":
            R = self.rproj()
            for q in set(R) | set(self.M[p + 1]):
                if q == _DONE:
                    X = np.tensordot(phi, R[q], axes=([3], [0]))             # [a, s, t, b']
                    Xs.append(X.reshape(a * s, t * c))
                    continue
                if q in R:
                    Rq = R[q]                                                # [b, x, b']
                else:
                    Rq = np.eye(c)[:, None, :]
                F = self.M[p + 1][q] if q in self.M[p + 1] else \
                    np.einsum("wu,ts->wtsu", np.eye(Rq.shape[1]), np.eye(t))
                T = np.einsum("astb,wrtx,bxc->aswrc", phi, F, Rq)           # [a, s, w, t', b']
                Xs.append(T.reshape(-1, t * c))
            return Xs
        raise ValueError(f"In noiseterm, got ortho = {ortho}, only supports `left` and `right`")

    def noiseterm(self, phi, ortho):
        Xs = self.noiseterm_factors(phi, ortho)
        if ortho == "left":
            return sum(X @ X.conj().T for X in Xs)
        return sum(X.T @ X.conj() for X in Xs)


# ---------------------------------------------------------------------------
# ProjMPOSum2
# ---------------------------------------------------------------------------

class ProjSum:
    """ProjMPOSum2 (projmposum2.jl:80-140): one ProjMPO per MPO; products and
    noise terms are summed; penalty projectors as in ProjMPOSum_MPS."""

    def __init__(self, mpos, dtype=np.float64, Ms=None, weight: float = 0.0):
        self.mpos = mpos
        self.parts = [O.Envs(W, dtype=dtype) for W in mpos]
        self.dtype = dtype
        self.Ms = Ms or []
        self.weight = weight
        self.m: List[np.ndarray] = []

    def fresh(self):
        return ProjSum(self.mpos, self.dtype, self.Ms, self.weight)

    @property
    def nsite(self):
        return self.parts[0].nsite

    @nsite.setter
    def nsite(self, n):
        for P in self.parts:
            P.nsite = n

    @property
    def lpos(self):
        return self.parts[0].lpos

    def position(self, A, pos):
        for P in self.parts:
            P.position(A, pos)
        self.m = [O.projector_local(A, M, pos, self.nsite) for M in self.Ms]

    def product(self, v):
        out = sum(P.product(v) for P in self.parts)
        for m in self.m:
            out = out + self.weight * np.vdot(m, v) * m
        return out

    def noiseterm(self, phi, ortho):
        return sum(O.noiseterm(P, phi, ortho) for P in self.parts)
