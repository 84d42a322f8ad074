"""FP64 CPU oracle for the tree-tensor-network (TTN) path -- TEST
INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the benchmark's CPU
baseline may import this module; the product path never does.

Restates, in numpy, the reference's TTN ground-state machinery:

* the default binary-tree graph and site placement
  (src/ttn/ttn_generators.jl:1-120), the default sweep path
  (src/ttn/sweep_ttn.jl:40-60) and the graph helpers (src/base/graph.jl:200-433);
* the random initial TTN and its isometrisation (ttn_generators.jl:200-279,
  src/ttn/ttn.jl:266-382);
* the effective Hamiltonian at a node and the environment moves between
  nodes (src/ttn/linktensors.jl:63-221, state_envs_ttn.jl:120-155);
* node updates, subspace expansion, sweeps and optimize!
  (src/ttn/update_site_ttn.jl:42-109, sweep_ttn.jl:117-274,
  optimize_ttn.jl:148-223).

The reference builds its environments as id-keyed sums of link tensors over
the terms of a CouplingModel.  This restatement sums the same terms through a
tree operator: every link carries channels (identity, completed subtree
Hamiltonian, and one open single-site operator per site operator whose
partner lies across the link) and every node a sparse coupling W between the
channels of its three legs.  The projected operator is the same linear map;
tests pin it against the dense Hamiltonian projected on the tree.

Parity unpinned (only running Julia could pin them): the tie-break between
equally central nodes (Julia Set order; here the lexicographically smallest
node), random initial tensors and subspace-expansion padding (Julia RNG; here
numpy default_rng), and the QR/SVD gauge of ITensors (any gauge gives the
same state).
"""

from __future__ import annotations

import math
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from oracle.tn_oracle import FLOAT64_THRESHOLD, eigsolve_lanczos, truncate_weights

Node = Tuple[int, int]

# ---------------------------------------------------------------------------
# Graph helpers (src/base/graph.jl).  Neighbour sets are iterated in sorted
# order so that every traversal is deterministic.
# ---------------------------------------------------------------------------


def min_pow2_ge(n: int) -> int:
    """_minimum_power2_greater_than (src/ttn/helper_internal_funcs.jl:10-20)."""
    p = 1
    while p < n:
        p <<= 1
    return p


def trailing_zeros(n: int) -> int:
    return (n & -n).bit_length() - 1


def distribute_site_positions(numsites: int, loc: Optional[List[int]] = None) -> List[int]:
    """_distribute_site_positions (ttn_generators.jl:1-37)."""
    if not loc:
        rem = numsites % 2
        loc = [numsites // 2, numsites // 2 + rem]
        numsites = min_pow2_ge(numsites)
    else:
        power = trailing_zeros(len(loc))
        new = []
        for m in loc:
            rem = m % 2
            if power % 2 == 1:
                new += [m // 2 + rem, m // 2]
            else:
                new += [m // 2, m // 2 + rem]
        loc = new
    numsites //= 2
    if trailing_zeros(numsites) > 1:
        return distribute_site_positions(numsites, loc)
    pos = []
    for m in loc:
        if m == 1:
            pos += [1, 0]
        elif m == 2:
            pos += [1, 1]
        else:
            raise RuntimeError("`_distribute_site_positions()`: SOMETHING IS WRONG !!")
    return pos


def default_graph_sitenodes(numsites: int):
    """default_graph_sitenodes (ttn_generators.jl:39-95): edges as a dict of
    neighbour sets and the site -> node map (sites 1-based)."""
    sp = distribute_site_positions(numsites)
    pow2 = len(sp)
    nlayers = trailing_zeros(pow2)
    edges: Dict[Node, set] = {}
    sitenodes: Dict[int, Node] = {}

    def addedge(a, b):
        edges.setdefault(a, set()).add(b)
        edges.setdefault(b, set()).add(a)

    count = 1
    for ll in range(1, nlayers - 1):
        for nn in range(1, (pow2 >> ll) + 1):
            if ll == 1 and (sp[2 * nn - 2] == 0 or sp[2 * nn - 1] == 0):
                sitenodes[count] = (ll + 1, (nn + 1) // 2)
                count += 1
                continue
            elif ll == 1:
                sitenodes[count] = (ll, nn)
                sitenodes[count + 1] = (ll, nn)
                count += 2
            addedge((ll, nn), (ll + 1, (nn + 1) // 2))
    addedge((nlayers - 1, 1), (nlayers - 1, 2))
    return edges, sitenodes


def bfs(edges, source):
    dist, parent = {source: 0}, {}
    q = deque([source])
    while q:
        n = q.popleft()
        for m in sorted(edges[n]):
            if m not in dist:
                dist[m] = dist[n] + 1
                parent[m] = n
                q.append(m)
    return dist, parent


def shortest_path(edges, a, b) -> List[Node]:
    if a == b:
        return [a]
    _, parent = bfs(edges, a)
    out = [b]
    while out[-1] != a:
        out.append(parent[out[-1]])
    return out[::-1]


def eccentric_center(edges, nodes=None) -> Node:
    """find_eccentric_central_node (graph.jl:400-433); ties -> smallest node."""
    best, bestd = None, None
    for n in sorted(edges):
        dist, _ = bfs(edges, n)
        d = max(v for k, v in dist.items() if nodes is None or k in nodes)
        if bestd is None or d < bestd:
            best, bestd = n, d
    return best


def default_sweeppath(numsites: int, edges) -> List[Node]:
    """default_sweeppath (sweep_ttn.jl:40-60)."""
    nsites = min_pow2_ge(numsites)
    nlayers = trailing_zeros(nsites)
    path = []
    for ll in range(nlayers - 1, 0, -1):
        cnt = nsites >> ll
        for nn in range(1, cnt + 1):
            nnpos = nn if (nlayers - ll) % 2 == 1 else cnt - nn + 1
            node = (ll, nnpos)
            if ll == 1 and node not in edges:
                continue
            path.append(node)
    return path


# ---------------------------------------------------------------------------
# Tree operator.  Rooted at `root`; every node has three slots: two children
# (node or site, ordered by their leftmost site) and the parent; the root's
# third slot is its third child.
# ---------------------------------------------------------------------------

SPIN_HALF = {
    "Id": np.eye(2),
    "Sz": np.diag([0.5, -0.5]),
    "S+": np.array([[0.0, 1.0], [0.0, 0.0]]),
    "S-": np.array([[0.0, 0.0], [1.0, 0.0]]),
    "Sx": np.array([[0.0, 0.5], [0.5, 0.0]]),
}


@dataclass
class Tree:
    N: int
    edges: Dict[Node, set]
    sitenodes: Dict[int, Node]
    root: Node
    parent: Dict[Node, Optional[Node]] = field(default_factory=dict)
    slots: Dict[Node, List[tuple]] = field(default_factory=dict)  # ("site", i) or ("node", m)
    sites_below: Dict[tuple, List[int]] = field(default_factory=dict)

    @staticmethod
    def build(N: int, edges=None, sitenodes=None, root=None) -> "Tree":
        if edges is None:
            edges, sitenodes = default_graph_sitenodes(N)
        if root is None:
            root = eccentric_center(edges, set(sitenodes.values()))
        t = Tree(N, edges, sitenodes, root)
        _, par = bfs(edges, root)
        t.parent = {n: par.get(n) for n in edges}
        kids: Dict[Node, List[tuple]] = {n: [] for n in edges}
        for i in range(1, N + 1):
            kids[sitenodes[i]].append(("site", i))
        for n in edges:
            if t.parent[n] is not None:
                kids[t.parent[n]].append(("node", n))

        def below(item) -> List[int]:
            if item in t.sites_below:
                return t.sites_below[item]
            if item[0] == "site":
                s = [item[1]]
            else:
                s = sorted(x for k in kids[item[1]] for x in below(k))
            t.sites_below[item] = s
            return s

        for n in edges:
            ks = sorted(kids[n], key=lambda k: below(k)[0])
            if t.parent[n] is None:
                if len(ks) != 3:
                    raise ValueError("tree root must have three legs")
                t.slots[n] = ks
            else:
                if len(ks) != 2:
                    raise ValueError("every non-root node must have two children and a parent")
                t.slots[n] = ks + [("node", t.parent[n])]
        for n in edges:
            below(("node", n))
        return t

    def slot_of(self, n: Node, m: Node) -> int:
        return self.slots[n].index(("node", m))

    def child_item(self, n: Node, s: int) -> tuple:
        """Item whose subtree is the far side of slot s, seen as a child
        side in the rooted tree (the slot's link channels belong to it)."""
        if s < 2 or self.parent[n] is None:
            return self.slots[n][s]
        return ("node", n)  # parent slot: the channels are those of n's own subtree


@dataclass
class CouplingTerms:
    """1- and 2-body spin terms: [(coef, [(opname, site), ...])], sites 1-based."""
    N: int
    terms: List[tuple]
    ops: Dict[str, np.ndarray] = field(default_factory=lambda: SPIN_HALF)


def build_tree_operator(tree: Tree, H: CouplingTerms):
    """Channels per link and sparse W per node (see the module docstring)."""
    onebody = {i: np.zeros((2, 2)) for i in range(1, tree.N + 1)}
    twobody: Dict[Tuple[int, str, int, str], float] = {}
    for coef, ops in H.terms:
        if len(ops) == 1:
            (o, i), = ops
            onebody[i] = onebody[i] + coef * H.ops[o]
        elif len(ops) == 2:
            (o1, i), (o2, j) = ops
            if i == j:
                onebody[i] = onebody[i] + coef * (H.ops[o1] @ H.ops[o2])
                continue
            if i > j:
                (o1, i), (o2, j) = (o2, j), (o1, i)
            key = (i, o1, j, o2)
            twobody[key] = twobody.get(key, 0.0) + coef
        else:
            raise ValueError("only 1- and 2-body terms are supported")
    partners: Dict[Tuple[int, str], List[Tuple[int, str, float]]] = {}
    for (i, o1, j, o2), J in twobody.items():
        partners.setdefault((i, o1), []).append((j, o2, J))
        partners.setdefault((j, o2), []).append((i, o1, J))

    def channels(item) -> List[tuple]:
        S = set(tree.sites_below[item])
        ch = [("I",), ("H",)]
        for i in sorted(S):
            for o in sorted({k[1] for k in partners if k[0] == i}):
                if any(j not in S for j, _, _ in partners[(i, o)]):
                    ch.append(("op", i, o))
        return ch

    chan = {item: channels(item) for item in tree.sites_below}
    W: Dict[Node, List[tuple]] = {}
    for n in tree.edges:
        items = [tree.child_item(n, s) for s in range(3)]
        idx = [{c: k for k, c in enumerate(chan[it])} for it in items]
        S = [set(tree.sites_below[it]) for it in items]
        terms: Dict[Tuple[int, int, int], float] = {}

        def add(c, w):
            terms[c] = terms.get(c, 0.0) + w

        I = [ix[("I",)] for ix in idx]
        Hh = [ix[("H",)] for ix in idx]
        if tree.parent[n] is None:
            add((Hh[0], I[1], I[2]), 1.0)
            add((I[0], Hh[1], I[2]), 1.0)
            add((I[0], I[1], Hh[2]), 1.0)
            for (i, o1, j, o2), J in twobody.items():
                a = [k for k in range(3) if i in S[k]][0]
                b = [k for k in range(3) if j in S[k]][0]
                if a == b:
                    continue
                c = list(I)
                c[a] = idx[a][("op", i, o1)]
                c[b] = idx[b][("op", j, o2)]
                add(tuple(c), J)
        else:
            add((I[0], I[1], I[2]), 1.0)
            add((Hh[0], I[1], Hh[2]), 1.0)
            add((I[0], Hh[1], Hh[2]), 1.0)
            for (i, o1, j, o2), J in twobody.items():
                a = [k for k in range(2) if i in S[k]]
                b = [k for k in range(2) if j in S[k]]
                if a and b and a[0] != b[0]:
                    c = [I[0], I[1], Hh[2]]
                    c[a[0]] = idx[a[0]][("op", i, o1)]
                    c[b[0]] = idx[b[0]][("op", j, o2)]
                    add(tuple(c), J)
            for k in range(2):
                for ch in chan[items[k]]:
                    if ch[0] == "op" and ch in idx[2]:
                        c = [I[0], I[1], idx[2][ch]]
                        c[k] = idx[k][ch]
                        add(tuple(c), 1.0)
        W[n] = sorted((c0, c1, c2, w) for (c0, c1, c2), w in terms.items())

    def site_env(i) -> np.ndarray:
        """Channel operators of a site leg, F[c][ket][bra] = <bra|X_c|ket>."""
        ch = chan[("site", i)]
        F = np.zeros((len(ch), 2, 2))
        for k, c in enumerate(ch):
            X = np.eye(2) if c[0] == "I" else onebody[i] if c[0] == "H" else H.ops[c[2]]
            F[k] = X.T
        return F

    return chan, W, {i: site_env(i) for i in range(1, tree.N + 1)}


# ---------------------------------------------------------------------------
# TTN state, environments and products.  Node tensors are numpy arrays with
# legs in slot order; environments F[c][ket][bra] per (node, slot), each
# describing the far side of that slot.
# ---------------------------------------------------------------------------


class TTNState:
    def __init__(self, tree: Tree, tensors: Dict[Node, np.ndarray], center: Optional[Node]):
        self.tree = tree
        self.T = tensors
        self.center = center

    def copy(self):
        return TTNState(self.tree, {k: v.copy() for k, v in self.T.items()}, self.center)

    def maxlinkdim(self) -> int:
        md = 1
        for n, A in self.T.items():
            for s, it in enumerate(self.tree.slots[n]):
                if it[0] == "node":
                    md = max(md, A.shape[s])
        return md


def _leg_first(A, s):
    return np.moveaxis(A, s, 0)


def move_isometry(psi: TTNState, n: Node, m: Node, maxdim=2 ** 62, mindim=1, cutoff=-1.0):
    """moveisometry_to_next! (ttn.jl:266-310): QR (or truncating SVD when
    cutoff > 0 or the bond exceeds maxdim) of node n with the leg toward m
    split off; the non-isometric factor is absorbed into m."""
    t = psi.tree
    s, sm = t.slot_of(n, m), t.slot_of(m, n)
    A = _leg_first(psi.T[n], s)
    shp = A.shape
    M = A.reshape(shp[0], -1).T  # rows: other legs, columns: the bond
    err = 0.0
    if cutoff > 0.0 or shp[0] > maxdim:
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        P = S * S
        tr = truncate_weights(P, maxdim, mindim, cutoff)
        k = tr.n
        err = tr.truncerr
        Q, R = U[:, :k], S[:k, None] * Vh[:k]
    else:
        Q, R = np.linalg.qr(M)
    k = Q.shape[1]
    psi.T[n] = np.moveaxis(Q.T.reshape((k,) + shp[1:]), 0, s)
    B = np.moveaxis(psi.T[m], sm, 0)
    bs = B.shape
    B = (R @ B.reshape(bs[0], -1)).reshape((k,) + bs[1:])
    psi.T[m] = np.moveaxis(B, 0, sm)
    psi.center = m
    return err


def isometrize_full(psi: TTNState, node: Node, **kw):
    """isometrize_full! (ttn.jl:312-340)."""
    dist, _ = bfs(psi.tree.edges, node)
    order = sorted(dist, key=lambda x: (-dist[x], x))
    for n in order[:-1]:
        nxt = shortest_path(psi.tree.edges, n, node)[1]
        move_isometry(psi, n, nxt, **kw)
    psi.center = node
    psi.T[node] = psi.T[node] / np.linalg.norm(psi.T[node])


def isometrize(psi: TTNState, node: Node, normalize=True, **kw):
    """isometrize! (ttn.jl:342-382)."""
    if psi.center is None:
        isometrize_full(psi, node, **kw)
        return
    path = shortest_path(psi.tree.edges, psi.center, node)
    for a, b in zip(path[:-1], path[1:]):
        move_isometry(psi, a, b, **kw)
    psi.center = node
    if normalize:
        psi.T[node] = psi.T[node] / np.linalg.norm(psi.T[node])


def random_ttn(tree: Tree, chi: int, seed: int) -> TTNState:
    """randomTTN / default_randomTTN for dense tensors (ttn_generators.jl:200-279):
    link dims min(prod of the node's other legs, chi), Gaussian entries (numpy
    default_rng(seed), nodes in sorted order), then isometrised on the root."""
    dims: Dict[tuple, int] = {}
    for i in range(1, tree.N + 1):
        dims[("site", i)] = 2
    center = tree.root
    dist, _ = bfs(tree.edges, center)
    for n in sorted(tree.edges, key=lambda x: (-dist[x], x)):
        if n == center:
            continue
        p = 1
        for it in tree.slots[n][:2]:
            p *= dims[it]
        dims[("node", n)] = min(p, chi)
    rng = np.random.default_rng(seed)
    T = {}
    for n in sorted(tree.edges):
        shape = tuple(dims[tree.child_item(n, s)] for s in range(3))
        A = rng.standard_normal(shape)
        T[n] = A / np.linalg.norm(A)
    psi = TTNState(tree, T, None)
    isometrize_full(psi, center, cutoff=0.0, maxdim=chi)
    return psi


class EnvsTTN:
    """Projected operator on a TTN (EnvCouplingModelTTN / LinkTensorsTTN,
    environment.jl:1-40, linktensors.jl:63-221)."""

    def __init__(self, psi: TTNState, H: CouplingTerms, Ms: Optional[List["TTNState"]] = None,
                 weight: float = 0.0):
        self.tree = psi.tree
        self.chan, self.W, site = build_tree_operator(psi.tree, H)
        # penalty projectors (EnvCouplingModelProjTTN, environment.jl:42-110;
        # LinkProjTTN, linkproj.jl): overlap environments O[(n, s)][x, y]
        # of the far side of slot s (psi index x, M index y; sites: identity)
        self.Ms = [M.T for M in (Ms or [])]
        self.weight = weight
        self.O: List[Dict[Tuple[Node, int], np.ndarray]] = [dict() for _ in self.Ms]
        self.F: Dict[Tuple[Node, int], np.ndarray] = {}
        for n in self.tree.edges:
            for s, it in enumerate(self.tree.slots[n]):
                if it[0] == "site":
                    self.F[(n, s)] = site[it[1]]
        # LinkTensorsTTN(psi, M): environments toward the center (linktensors.jl:240-262)
        c = psi.center
        dist, _ = bfs(self.tree.edges, c)
        for n in sorted(self.tree.edges, key=lambda x: (-dist[x], x)):
            if n != c:
                self.update(psi, n, shortest_path(self.tree.edges, n, c)[1])

    def update(self, psi: TTNState, n: Node, m: Node):
        """move_linktensors_to_next!(env, psi, n, m) (linktensors.jl:63-118):
        the environment of the link n -> m from n's side."""
        t = self.tree
        s = t.slot_of(n, m)
        v = psi.T[n]
        others = [k for k in range(3) if k != s]
        nc = len(self.chan[t.child_item(n, s)])
        out = np.zeros((nc, v.shape[s], v.shape[s]))
        for c0, c1, c2, w in self.W[n]:
            c = (c0, c1, c2)
            Fa, Fb = self.F[(n, others[0])][c[others[0]]], self.F[(n, others[1])][c[others[1]]]
            if s == 2:
                out[c2] += w * np.einsum("ax,by,abk,xyl->kl", Fa, Fb, v, v, optimize=True)
            elif s == 1:
                out[c1] += w * np.einsum("ax,cz,akc,xlz->kl", Fa, Fb, v, v, optimize=True)
            else:
                out[c0] += w * np.einsum("by,cz,kbc,lyz->kl", Fa, Fb, v, v, optimize=True)
        self.F[(m, t.slot_of(m, n))] = out
        for M, O in zip(self.Ms, self.O):
            # move_linkproj_to_next! (linkproj.jl:48-66): conj(M[n]) x psi[n] x the other links
            T = v
            for o in others:
                if (n, o) in O:
                    T = np.moveaxis(np.tensordot(T, O[(n, o)], axes=([o], [0])), -1, o)
            Mn = M[n].conj()
            O[(m, t.slot_of(m, n))] = np.tensordot(T, Mn, axes=(others, others))

    def local_projection(self, n: Node, j: int) -> np.ndarray:
        """m[x] = sum_y M[y] prod_s conj(O_s[x_s, y_s]) at node n (contract /
        product of LinkProjTTN, linkproj.jl:126-165): P v = m <m, v>."""
        T = self.Ms[j][n]
        for s in range(3):
            if (n, s) in self.O[j]:
                T = np.moveaxis(np.tensordot(T, self.O[j][(n, s)].conj(), axes=([s], [1])), -1, s)
        return T

    def move(self, psi: TTNState, src: Node, dst: Node, skip: Optional[Node] = None):
        """move_linktensors! (linktensors.jl:120-150)."""
        path = shortest_path(self.tree.edges, src, dst)
        for a, b in zip(path[:-1], path[1:]):
            if a == skip:
                continue
            self.update(psi, a, b)

    def product(self, n: Node, v: np.ndarray) -> np.ndarray:
        """product(LinkTensorsTTN, psi, v) at the center n (linktensors.jl:183-221)."""
        out = np.zeros_like(v)
        for c0, c1, c2, w in self.W[n]:
            out += w * np.einsum("ax,by,cz,abc->xyz", self.F[(n, 0)][c0], self.F[(n, 1)][c1],
                                 self.F[(n, 2)][c2], v, optimize=True)
        for j in range(len(self.Ms)):
            m = self.local_projection(n, j)
            out = out + self.weight * np.vdot(m, v) * m
        return out


class StateEnvsTTN:
    """StateEnvsTTN(psi, H) (state_envs_ttn.jl:20-30) or StateEnvsTTN(psi, H,
    Ms; weight) (state_envs_ttn.jl:36-70): copies psi."""

    def __init__(self, psi: TTNState, H: CouplingTerms, Ms: Optional[List[TTNState]] = None,
                 weight: Optional[float] = None):
        if Ms is not None and (weight is None or weight <= 0.0):
            raise ValueError(f"`weight` parameter should be > 0.0 (value passed was `weight={weight}`)")
        self.psi = psi.copy()
        self.env = EnvsTTN(self.psi, H, Ms, weight or 0.0)

    def position(self, node: Node, skip: Optional[Node] = None, normalize=True, **kw):
        """position!(sysenv, node; ...) (state_envs_ttn.jl:120-140)."""
        oc = self.psi.center
        isometrize(self.psi, node, normalize=normalize, **kw)
        self.env.move(self.psi, oc, node, skip)


def update_position(sysenv: StateEnvsTTN, node: Node, *, maxdim=2 ** 62, mindim=1, cutoff=-1.0,
                    normalize=True, krylovdim=5, maxiter=2, tol=1e-14):
    """update_position! (update_site_ttn.jl:42-63) with eig_solver."""
    sysenv.position(node, normalize=normalize, maxdim=maxdim, mindim=mindim, cutoff=cutoff)
    v = sysenv.psi.T[node]
    e, v, _ = eigsolve_lanczos(lambda x: sysenv.env.product(node, x), v, krylovdim, maxiter, tol)
    sysenv.psi.T[node] = v
    return e


def subspace_expand(psi: TTNState, node: Node, nxt: Node, max_expand_dim: int, noise: float,
                    rng: np.random.Generator):
    """subspace_expand! (update_site_ttn.jl:65-109) for dense tensors: both
    tensors grow along their common bond by p = min(sum of the other leg
    dims of each, max_expand_dim) with Gaussian padding scaled to
    noise * |tensor| (padding for nxt drawn first, then for node)."""
    t = psi.tree
    sa, sb = t.slot_of(node, nxt), t.slot_of(nxt, node)
    A, B = psi.T[node], psi.T[nxt]
    sumA = sum(A.shape[k] for k in range(3) if k != sa)
    sumB = sum(B.shape[k] for k in range(3) if k != sb)
    p = min(sumA, sumB, max_expand_dim)

    def pad(X, s):
        shape = list(X.shape)
        shape[s] = p
        P = rng.standard_normal(shape)
        P *= noise * np.linalg.norm(X) / np.linalg.norm(P)
        return np.concatenate([X, P], axis=s)

    psi.T[nxt] = pad(B, sb)
    psi.T[node] = pad(A, sa)


@dataclass
class SweepDataTTN:
    """sweep_ttn.jl:1-30."""
    sweepcount: int = 0
    maxchi: List[int] = field(default_factory=list)
    energy: List[float] = field(default_factory=list)


def fullsweep(sysenv: StateEnvsTTN, path: List[Node], swdata: SweepDataTTN, *, maxdim=2 ** 62, mindim=1,
              cutoff=FLOAT64_THRESHOLD, noise=0.0, expand_dim=None, max_expand_dim=None, expand_numiter=4,
              krylovdim=5, maxiter=2, tol=1e-14, rng=None):
    """fullsweep!(sysenv, sweeppath, eig_solver, swdata) (sweep_ttn.jl:117-274)."""
    if expand_dim is None:
        expand_dim = 0 if abs(noise) < 100 * FLOAT64_THRESHOLD else 20
    if max_expand_dim is None:
        max_expand_dim = 2 * expand_dim
    if expand_dim != 0 and expand_numiter < 2:
        raise ValueError(f"`fullsweep!()`: `expand_numiter={expand_numiter}` cannot be less than 2 "
                         f"for `expand_dim={expand_dim}` !!")
    kw = dict(krylovdim=krylovdim, maxiter=maxiter, tol=tol)
    energy = float("nan")
    swdata.sweepcount += 1
    if abs(noise) < 100 * FLOAT64_THRESHOLD:
        for node in path + path[::-1]:
            energy = update_position(sysenv, node, maxdim=maxdim, mindim=mindim, cutoff=-1.0, **kw)
    else:
        rng = rng if rng is not None else np.random.default_rng(0)
        central = eccentric_center(sysenv.psi.tree.edges)
        n = len(path)
        for ii in list(range(n)) + list(range(n - 1, -1, -1)):
            node = path[ii]
            if node == central:
                energy = update_position(sysenv, node, maxdim=maxdim, mindim=mindim, cutoff=-1.0, **kw)
                continue
            nxt = shortest_path(sysenv.psi.tree.edges, node, central)[1]
            sysenv.position(node, skip=nxt, maxdim=maxdim, mindim=mindim, cutoff=-1.0)
            subspace_expand(sysenv.psi, node, nxt, max_expand_dim, noise, rng)
            sysenv.psi.center = nxt
            for dummy in range(1, expand_numiter + 1):
                if dummy == expand_numiter:
                    md = maxdim
                elif dummy == 1:
                    md = maxdim + max_expand_dim
                else:
                    md = maxdim + expand_dim
                newnode = node if dummy % 2 == 1 else nxt
                energy = update_position(sysenv, newnode, maxdim=md, mindim=mindim, cutoff=cutoff, **kw)
    swdata.maxchi.append(sysenv.psi.maxlinkdim())
    swdata.energy.append(energy)
    if swdata.sweepcount > 1:
        return swdata.energy[-1] - swdata.energy[-2]
    return float("nan")


@dataclass
class OptimizeParamsTTN:
    """OptimizeParamsTTN (optimize_ttn.jl:1-100)."""
    maxdim: List[int]
    nsweeps: List[int]
    cutoff: List[float]
    noise: List[float]
    noisedecay: List[float]
    disable_noise_after: List[int]

    @staticmethod
    def make(maxdim, nsweeps, cutoff=FLOAT64_THRESHOLD, noise=0.0, noisedecay=1.0,
             disable_noise_after=2 ** 62):
        n = len(nsweeps)
        f = lambda x, cast: list(x) if isinstance(x, (list, tuple)) else [cast(x)] * n
        p = OptimizeParamsTTN(list(maxdim), list(nsweeps), f(cutoff, float), f(noise, float),
                              f(noisedecay, float), f(disable_noise_after, int))
        if not (len(p.nsweeps) == len(p.maxdim) == len(p.cutoff) == len(p.noise) == len(p.noisedecay)
                == len(p.disable_noise_after)):
            raise ValueError("`OptimizeParamsTTN()` :: Size mismatch in input vectors !!")
        return p


def optimize(sysenv: StateEnvsTTN, params: OptimizeParamsTTN, path: List[Node], *, energyErrGoal=None,
             seed=0, **kw) -> SweepDataTTN:
    """optimize!(sysenv, params, sweeppath) (optimize_ttn.jl:148-223)."""
    sw = SweepDataTTN()
    rng = np.random.default_rng(seed)
    for ii in range(len(params.nsweeps)):
        noise = params.noise[ii]
        for jj in range(1, params.nsweeps[ii] + 1):
            enerr = fullsweep(sysenv, path, sw, maxdim=params.maxdim[ii], cutoff=params.cutoff[ii],
                              noise=noise, rng=rng, **kw)
            if energyErrGoal is not None and abs(enerr) < abs(energyErrGoal) and abs(noise) < FLOAT64_THRESHOLD:
                break
            if jj == params.disable_noise_after[ii]:
                noise = 0.0
            noise /= params.noisedecay[ii]
            if noise < 100 * FLOAT64_THRESHOLD:
                noise = 0.0
    return sw


def heisenberg_terms(N: int, J: float = 1.0, h: Optional[Sequence[float]] = None) -> CouplingTerms:
    """sum_j J (Sz Sz + (S+S- + S-S+)/2) + sum_j h_j Sz_j in OpStrings form
    (the reference's test model, test/test_TTN.jl:6-24, plus fields)."""
    terms = []
    for j in range(1, N):
        terms.append((J, [("Sz", j), ("Sz", j + 1)]))
        terms.append((0.5 * J, [("S+", j), ("S-", j + 1)]))
        terms.append((0.5 * J, [("S-", j), ("S+", j + 1)]))
    if h is not None:
        for j in range(1, N + 1):
            terms.append((float(h[j - 1]), [("Sz", j)]))
    return CouplingTerms(N, terms)
