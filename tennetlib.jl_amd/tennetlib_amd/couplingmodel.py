"""OpStrings and CouplingModel (src/base/opstrings.jl, src/base/couplingmodel.jl)
and their lowering to an MPO that the StateEnvs path runs on the GPU.

A CouplingModel is a list of term ids.  Each id is an operator string on a
set of sites (its support) whose factors are joined by small virtual links
(the reference's "OpLink"s): all OpStrings terms on one support are merged
into one id and compressed by two SVD sweeps, every single-site term of a
site is summed into one id, and every MPO handed to the constructor is one
id on all sites.  ProjCouplingModel (src/mps/projcouplingmodel.jl) contracts
each id separately; its product is H_eff and its noise term is the sum over
ids of X_id X_id^H, where the id's environment and local factors act on phi,
over the ids that have already acted at or left of the bond (left) or at or
right of it (right).

``coupling_mpo`` writes the same sum as one MPO whose bond channels are, in
this order: the identity string with all of its terms still to the right
(split into one channel per id that starts at the next site plus one for the
rest), one block per id open across the bond (its virtual link), and the
completed terms (one channel per id that ended at this site plus one for the
rest).  The product of this MPO is the product of ProjCouplingModel, and its
noise factor restricted to the channels that carry an operator on the
kept side (the leading identity-on-the-left channels dropped for the left
factor, the trailing identity-on-the-right channels for the right factor) is
the concatenation of the per-id factors, so X X^H is the same sum.
"""

from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

FLOAT64_THRESHOLD = 1e-15
CHUNKSIZE = 12

SPIN_HALF_OPS = {
    "Id": np.eye(2), "Sz": np.diag([0.5, -0.5]),
    "S+": np.array([[0.0, 1.0], [0.0, 0.0]]), "S-": np.array([[0.0, 0.0], [1.0, 0.0]]),
    "Sx": np.array([[0.0, 0.5], [0.5, 0.0]]),
}


def op_matrix(name: str) -> np.ndarray:
    """op(name, site) for spin-1/2 sites; "A * B" is the matrix product A B
    (the form bosonize gives several operators on one site, fermions.jl:79-91)."""
    M = np.eye(2)
    for part in name.split(" * "):
        part = part.strip()
        if part not in SPIN_HALF_OPS:
            raise ValueError(f"CouplingModel: unknown operator {part}")
        M = M @ SPIN_HALF_OPS[part]
    return M


class OpStrings:
    """OpStrings (src/base/opstrings.jl): ``os += (coef, "Sz", j, "Sz", j + 1)``
    or ``os += (coef, ("Sz", j), ("Sz", j + 1))``."""

    def __init__(self):
        self.terms: List[Tuple[float, List[Tuple[str, int]]]] = []

    def __iadd__(self, t):
        coef, rest = float(t[0]), list(t[1:])
        if rest and isinstance(rest[0], tuple):
            ops = [(str(o), int(i)) for o, i in rest]
        else:
            if len(rest) % 2:
                raise ValueError("OpStrings: operator names and sites must come in pairs")
            ops = [(str(rest[k]), int(rest[k + 1])) for k in range(0, len(rest), 2)]
        self.terms.append((coef, ops))
        return self


def _normalise_terms(os: OpStrings, N: int) -> List[Tuple[float, Tuple[Tuple[str, int], ...]]]:
    """removeIdsZeros, bosonize and mergeterms (opstrings.jl:161-237,
    fermions.jl:30-96) for spin-1/2 operators: "Id" factors dropped, terms
    with |coef| <= 1e-15 dropped, operators on one site multiplied in the
    order given, sites sorted, identical strings summed.  Strings whose sum
    cancels are dropped too (the reference would compress them to 0/0)."""
    merged: Dict[Tuple[Tuple[str, int], ...], float] = {}
    for coef, ops in os.terms:
        ops = [(o, i) for o, i in ops if o != "Id"]
        if abs(coef) <= FLOAT64_THRESHOLD:
            continue
        if not ops:
            raise ValueError("CouplingModel: a term made of identity operators only")
        for o, i in ops:
            op_matrix(o)
            if not 1 <= i <= N:
                raise ValueError("`CouplingModel()`: Out-of-bounds error in OpStrings !!")
        by_site: Dict[int, str] = {}
        for o, i in ops:
            by_site[i] = f"{by_site[i]} * {o}" if i in by_site else o
        key = tuple((by_site[i], i) for i in sorted(by_site))
        merged[key] = merged.get(key, 0.0) + coef
    out = [(c, k) for k, c in merged.items() if abs(c) > FLOAT64_THRESHOLD]
    out.sort(key=lambda t: (t[1][0][1], len(t[1])))
    return out


def _term_factors(coef: float, ops) -> List[np.ndarray]:
    """One string as factors F[l, s', s, r] with unit links: every factor
    scaled by |coef|^(1/n), the sign on the first (couplingmodel.jl:175-185)."""
    n = len(ops)
    scale = abs(coef) ** (1.0 / n)
    F = [(scale * op_matrix(o))[None, :, :, None] for o, _ in ops]
    F[0] = F[0] * np.sign(coef)
    return F


def _directsum(strings: List[List[np.ndarray]]) -> List[np.ndarray]:
    """Direct sum of strings on one support along their links (first factor
    concatenated along its right link, last along its left, others block
    diagonal)."""
    n = len(strings[0])
    out = []
    for j in range(n):
        ls = [s[j].shape[0] for s in strings]
        rs = [s[j].shape[3] for s in strings]
        d1, d2 = strings[0][j].shape[1:3]
        Lt = 1 if j == 0 else sum(ls)
        Rt = 1 if j == n - 1 else sum(rs)
        T = np.zeros((Lt, d1, d2, Rt), dtype=np.result_type(*[s[j] for s in strings]))
        lo = ro = 0
        for s, l, r in zip(strings, ls, rs):
            li = slice(0, 1) if j == 0 else slice(lo, lo + l)
            ri = slice(0, 1) if j == n - 1 else slice(ro, ro + r)
            T[li, :, :, ri] += s[j]
            lo += l
            ro += r
        out.append(T)
    return out


def _truncate_count(P: np.ndarray, cutoff: float) -> int:
    """Number of values kept by ITensors' svd truncation with maxdim =
    typemax, mindim = 1 and a relative cutoff (discarded weight <= cutoff * sum)."""
    total = float(P.sum())
    n = P.size
    err = 0.0
    while n > 1 and err + P[n - 1] <= cutoff * total:
        err += P[n - 1]
        n -= 1
    return n


def _compress(strings: List[List[np.ndarray]]) -> List[np.ndarray]:
    """_chunksum_cm_terms for at most CHUNKSIZE strings
    (couplingmodel.jl:65-104): direct sum, then a left and a right sweep of
    SVDs; at every bond the singular values are normalised and absorbed
    toward the sweep direction and both factors carry sqrt(|S|)."""
    T = _directsum(strings)
    n = len(T)
    for direction in ("left", "right"):
        bonds = range(n - 1) if direction == "left" else range(n - 2, -1, -1)
        for b in bonds:
            l, d1, d2, _ = T[b].shape
            _, e1, e2, r = T[b + 1].shape
            phi = np.tensordot(T[b], T[b + 1], axes=([3], [0])).reshape(l * d1 * d2, e1 * e2 * r)
            U, S, Vh = np.linalg.svd(phi, full_matrices=False)
            k = _truncate_count(S * S, FLOAT64_THRESHOLD)
            U, S, Vh = U[:, :k], S[:k], Vh[:k]
            norm = float(np.linalg.norm(S))
            S = S / norm
            if direction == "left":
                Vh = S[:, None] * Vh
            else:
                U = U * S[None, :]
            T[b] = (U * np.sqrt(norm)).reshape(l, d1, d2, k)
            T[b + 1] = (Vh * np.sqrt(norm)).reshape(k, e1, e2, r)
    return T


def _chunksum(strings: List[List[np.ndarray]]) -> List[np.ndarray]:
    """Sum of the strings of one support, in chunks of CHUNKSIZE; the chunk
    sums are summed again (_divide_by_chunksize, helper_internal_funcs.jl:21-32)."""
    if len(strings) == 1:
        return strings[0]
    if len(strings) <= CHUNKSIZE:
        return _compress(strings)
    nfull, left = divmod(len(strings), CHUNKSIZE)
    chunks = [strings[i * CHUNKSIZE:(i + 1) * CHUNKSIZE] for i in range(nfull)]
    if left:
        chunks.append(strings[nfull * CHUNKSIZE:])
    return _chunksum([_chunksum(c) for c in chunks])


Id = Tuple[Tuple[int, ...], List[np.ndarray]]


class CouplingModel:
    """CouplingModel (src/base/couplingmodel.jl):

    * ``CouplingModel(os, sites)`` -- OpStrings on ``sites`` spin-1/2 sites
      (an int or a sequence of that length), terms on one support merged
      (``merge=True``, couplingmodel.jl:153-241);
    * ``CouplingModel(os, mpo)`` -- the OpStrings plus the MPO as one more
      term (couplingmodel.jl:292-322);
    * ``CouplingModel(mpo1, mpo2, ...)`` -- each MPO one term
      (couplingmodel.jl:331-356).

    MPOs are lists of W[wl, s', s, wr] with unit outer links."""

    def __init__(self, *args, merge: bool = True):
        self.mpos: List[List[np.ndarray]] = []
        self.terms: List[Tuple[float, Tuple[Tuple[str, int], ...]]] = []
        self.merge = merge
        if args and isinstance(args[0], OpStrings):
            if len(args) != 2:
                raise TypeError("CouplingModel(os, sites) or CouplingModel(os, mpo)")
            os, second = args
            if isinstance(second, (int, np.integer)):
                self.N = int(second)
            elif len(second) and isinstance(second[0], np.ndarray) and second[0].ndim == 4:
                self.mpos.append([np.asarray(W, dtype=np.float64) for W in second])
                self.N = len(second)
            else:
                self.N = len(second)
            self.terms = _normalise_terms(os, self.N)
        else:
            if not args:
                raise TypeError("CouplingModel needs OpStrings and sites, or MPOs")
            self.mpos = [[np.asarray(W, dtype=np.float64) for W in m] for m in args]
            self.N = len(self.mpos[0])
            d0 = [W.shape[1] for W in self.mpos[0]]
            if any([W.shape[1] for W in m] != d0 for m in self.mpos):
                raise ValueError("CouplingModel: the MPOs act on different sites")
        for m in self.mpos:
            if len(m) != self.N or m[0].shape[0] != 1 or m[-1].shape[3] != 1:
                raise ValueError("CouplingModel: an MPO must have N sites and unit outer links")
        self.d = [int(W.shape[1]) for W in self.mpos[0]] if self.mpos else [2] * self.N
        if self.mpos and self.terms and any(d != 2 for d in self.d):
            raise ValueError("CouplingModel: OpStrings act on spin-1/2 sites, the MPO does not")

    def __len__(self):
        return self.N

    # -- the 1- and 2-body view used by the tree-operator builder ------------

    def _few_body(self):
        if self.mpos:
            raise NotImplementedError("CouplingModel with MPO terms on the TTN path")
        one = {i: np.zeros((2, 2)) for i in range(1, self.N + 1)}
        two: Dict[Tuple[int, str, int, str], float] = {}
        for coef, ops in self.terms:
            if len(ops) == 1:
                (o, i), = ops
                one[i] = one[i] + coef * op_matrix(o)
            elif len(ops) == 2:
                (o1, i), (o2, j) = ops
                two[(i, o1, j, o2)] = two.get((i, o1, j, o2), 0.0) + coef
            else:
                raise NotImplementedError("CouplingModel on the TTN path: 1- and 2-body terms only")
        return one, two

    @property
    def onebody(self) -> Dict[int, np.ndarray]:
        return self._few_body()[0]

    @property
    def twobody(self) -> Dict[Tuple[int, str, int, str], float]:
        return self._few_body()[1]

    @property
    def partners(self) -> Dict[Tuple[int, str], List[int]]:
        p: Dict[Tuple[int, str], List[int]] = {}
        for (i, o1, j, o2) in self.twobody:
            p.setdefault((i, o1), []).append(j)
            p.setdefault((j, o2), []).append(i)
        return p

    # -- term ids (couplingmodel.jl:153-241) --------------------------------

    def ids(self) -> List[Id]:
        """(support, factors) of every term id; supports are 1-based sites."""
        groups: Dict[Tuple[int, ...], List[List[np.ndarray]]] = {}
        for coef, ops in self.terms:
            pos = tuple(i for _, i in ops)
            F = _term_factors(coef, ops)
            if pos in groups and len(pos) == 1:
                groups[pos][0] = [groups[pos][0][0] + F[0]]
            else:
                groups.setdefault(pos, []).append(F)
        out: List[Id] = []
        for pos, strings in groups.items():
            if self.merge:
                out.append((pos, _chunksum(strings)))
            else:
                out.extend((pos, s) for s in strings)
        full = tuple(range(1, self.N + 1))
        out.extend((full, list(m)) for m in self.mpos)
        return out


def mpo_sum_ids(mpos: Sequence[Sequence[np.ndarray]]) -> List[Id]:
    """Each MPO of a sum as one id on all sites (ProjMPOSum2, projmposum2.jl)."""
    N = len(mpos[0])
    return [(tuple(range(1, N + 1)), [np.asarray(W, dtype=np.float64) for W in m]) for m in mpos]


def coupling_mpo(ids: Sequence[Id], N: int, d: Sequence[int]):
    """The ids as one MPO W[wl, s', s, wr] with unit outer links, and per bond
    k (0..N) the number of leading channels with the identity on sites <= k
    (``skip_left``) and of trailing channels with the identity on sites > k
    (``skip_right``).  Channel blocks of bond k:

      ("IL",)   identity on 1..k, then every id with all sites >= k+2
      ("S", q)  identity on 1..k, then id q, which starts at k+1
      ("O", q)  id q open across the bond (its virtual link)
      ("E", q)  id q, which ended at k, then the identity
      ("IR",)   every id with all sites <= k-1, then the identity
    """
    sup = [tuple(p) for p, _ in ids]
    for p, F in ids:
        if list(p) != sorted(set(p)) or not (1 <= p[0] and p[-1] <= N) or len(F) != len(p):
            raise ValueError("coupling_mpo: an id needs sorted distinct sites in 1..N and one factor per site")
        if F[0].shape[0] != 1 or F[-1].shape[3] != 1:
            raise ValueError("coupling_mpo: the outer links of an id must have dimension 1")
        for q, i in enumerate(p):
            if F[q].shape[1] != d[i - 1] or F[q].shape[2] != d[i - 1]:
                raise ValueError(f"coupling_mpo: factor of a term does not match site {i}")
            if q + 1 < len(p) and F[q].shape[3] != F[q + 1].shape[0]:
                raise ValueError("coupling_mpo: virtual links of a term do not match")
    nid = len(ids)

    def link_dim(q, k):  # dimension of id q's link across bond k (min <= k < max)
        p, F = ids[q]
        last = max(t for t, i in enumerate(p) if i <= k)
        return F[last].shape[3]

    blocks: List[List[Tuple[tuple, int]]] = []
    skip_l, skip_r = [0] * (N + 1), [0] * (N + 1)
    for k in range(N + 1):
        bl: List[Tuple[tuple, int]] = []
        if k + 2 <= N and any(p[0] >= k + 2 for p in sup):
            bl.append((("IL",), 1))
        bl += [(("S", q), 1) for q in range(nid) if sup[q][0] == k + 1]
        skip_l[k] = len(bl)
        bl += [(("O", q), link_dim(q, k)) for q in range(nid) if sup[q][0] <= k < sup[q][-1]]
        tail = [(("E", q), 1) for q in range(nid) if sup[q][-1] == k]
        if k >= 2 and any(p[-1] <= k - 1 for p in sup):
            tail.append((("IR",), 1))
        skip_r[k] = len(tail)
        blocks.append(bl + tail)
    offs = []
    for bl in blocks:
        o, m = 0, {}
        for key, dim in bl:
            m[key] = (o, dim)
            o += dim
        offs.append((m, o))
    Ws = []
    for k in range(1, N + 1):
        (lm, wl), (rm, wr) = offs[k - 1], offs[k]
        dk = d[k - 1]
        W = np.zeros((wl, dk, dk, wr))
        I = np.eye(dk)

        def put(lkey, rkey, T):  # T[l, s', s, r]
            lo, ld = lm[lkey]
            ro, rd = rm[rkey]
            W[lo:lo + ld, :, :, ro:ro + rd] += T

        if ("IL",) in lm:
            if ("IL",) in rm:
                put(("IL",), ("IL",), I[None, :, :, None])
            for q in range(nid):
                if sup[q][0] == k + 1:
                    put(("IL",), ("S", q), I[None, :, :, None])
        for q in range(nid):
            p, F = ids[q]
            if sup[q][0] == k:  # id q starts here
                put(("S", q), ("E", q) if p[-1] == k else ("O", q), F[0])
            elif p[0] < k <= p[-1]:
                if k in p:
                    put(("O", q), ("E", q) if p[-1] == k else ("O", q), F[p.index(k)])
                else:
                    ld = lm[("O", q)][1]
                    put(("O", q), ("O", q), np.einsum("ab,st->astb", np.eye(ld), I))
            elif p[-1] == k - 1:
                put(("E", q), ("IR",), I[None, :, :, None])
        if ("IR",) in lm:
            put(("IR",), ("IR",), I[None, :, :, None])
        Ws.append(W)
    # unit outer links: the first tensor sums its rows, the last its columns
    Ws[0] = Ws[0].sum(axis=0, keepdims=True)
    Ws[-1] = Ws[-1].sum(axis=3, keepdims=True)
    skip_l[0] = skip_r[0] = skip_l[N] = skip_r[N] = 0
    return Ws, skip_l, skip_r
