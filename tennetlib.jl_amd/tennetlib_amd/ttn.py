"""Tree tensor networks on the GPU: host mirror of TenNetLib.jl's TTN API
(src/ttn/*.jl) over the tn_ttn C ABI of libtnhip.so.

Mirrors: default_graph_sitenodes / default_randomTTN (ttn_generators.jl),
default_sweeppath (sweep_ttn.jl:40-60), TTN (ttn.jl), StateEnvsTTN and
position! (state_envs_ttn.jl), update_position! / subspace_expand!
(update_site_ttn.jl), fullsweep! / SweepDataTTN (sweep_ttn.jl),
OptimizeParamsTTN / optimize! / optimize (optimize_ttn.jl), and the
CouplingModel built from OpStrings (src/base/couplingmodel.jl,
src/base/opstrings.jl) for 1- and 2-body spin-1/2 terms.

The Hamiltonian reaches the GPU as a tree operator: every link carries the
channels (identity, completed subtree Hamiltonian, one open single-site
operator per site operator with a partner across the link) and every node a
sparse coupling between the channels of its three legs.  Node tensors,
environments, products, Lanczos updates and QR/SVD isometry moves run on the
device; this module keeps the graph, the paths and the sweep logic.
"""

from __future__ import annotations

import ctypes as C
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .core import FLOAT64_THRESHOLD, Context, _live_children, default_context, eig_solver
from .couplingmodel import SPIN_HALF_OPS, CouplingModel, OpStrings, op_matrix  # noqa: F401

Node = Tuple[int, int]
TYPEMAX_INT = 2 ** 63 - 1

# ---------------------------------------------------------------------------
# graph helpers (src/base/graph.jl, src/ttn/ttn_generators.jl); neighbour
# sets are walked in sorted order so that every traversal is deterministic
# ---------------------------------------------------------------------------


def _pow2_ge(n: int) -> int:
    """_minimum_power2_greater_than (src/ttn/helper_internal_funcs.jl:10-20)."""
    if n <= 0:
        raise ValueError("`_minimum_power2_greater_than()`: Input must be a positive integer")
    return 1 << (n - 1).bit_length()


def _tz(n: int) -> int:
    return (n & -n).bit_length() - 1


def _site_positions(numsites: int, loc: Optional[List[int]] = None) -> List[int]:
    """_distribute_site_positions (ttn_generators.jl:1-37)."""
    if not loc:
        loc = [numsites // 2, numsites // 2 + numsites % 2]
        numsites = _pow2_ge(numsites)
    else:
        odd = _tz(len(loc)) % 2 == 1
        loc = [x for m in loc for x in ((m // 2 + m % 2, m // 2) if odd else (m // 2, m // 2 + m % 2))]
    numsites //= 2
    if _tz(numsites) > 1:
        return _site_positions(numsites, loc)
    out = []
    for m in loc:
        if m not in (1, 2):
            raise RuntimeError("`_distribute_site_positions()`: SOMETHING IS WRONG !!")
        out += [1, 0] if m == 1 else [1, 1]
    return out


def default_graph_sitenodes(numsites: int):
    """default_graph_sitenodes(N) (ttn_generators.jl:39-95) -> (graph, sitenodes):
    graph as {node: set(neighbours)}, sitenodes as {site (1-based): node}."""
    if numsites < 5:
        raise ValueError("default_graph_sitenodes: needs at least 5 sites")
    sp = _site_positions(numsites)
    nlayers = _tz(len(sp))
    graph: Dict[Node, set] = {}
    sitenodes: Dict[int, Node] = {}

    def edge(a, b):
        graph.setdefault(a, set()).add(b)
        graph.setdefault(b, set()).add(a)

    site = 1
    for ll in range(1, nlayers - 1):
        for nn in range(1, (len(sp) >> ll) + 1):
            if ll == 1:
                if sp[2 * nn - 2] == 0 or sp[2 * nn - 1] == 0:
                    sitenodes[site] = (2, (nn + 1) // 2)
                    site += 1
                    continue
                sitenodes[site] = sitenodes[site + 1] = (1, nn)
                site += 2
            edge((ll, nn), (ll + 1, (nn + 1) // 2))
    edge((nlayers - 1, 1), (nlayers - 1, 2))
    return graph, sitenodes


def _bfs(graph, src):
    dist, par = {src: 0}, {}
    q = deque([src])
    while q:
        n = q.popleft()
        for m in sorted(graph[n]):
            if m not in dist:
                dist[m], par[m] = dist[n] + 1, n
                q.append(m)
    return dist, par


def shortest_path(graph, a: Node, b: Node) -> List[Node]:
    """shortest_path (graph.jl:286-305)."""
    if a == b:
        return [a]
    _, par = _bfs(graph, a)
    if b not in par:
        raise ValueError(f"`shortest_path()`: `destination={b}` is not reachable from `source={a}` !!")
    out = [b]
    while out[-1] != a:
        out.append(par[out[-1]])
    return out[::-1]


def nextnode_in_path(graph, a: Node, b: Node, n: int = 1) -> Node:
    return shortest_path(graph, a, b)[n]


def find_eccentric_central_node(graph, nodes=None) -> Node:
    """find_eccentric_central_node (graph.jl:400-433); among equally central
    nodes the smallest (layer, index) is taken."""
    best = None
    for n in sorted(graph):
        dist, _ = _bfs(graph, n)
        e = max(d for k, d in dist.items() if nodes is None or k in nodes)
        if best is None or e < best[0]:
            best = (e, n)
    return best[1]


# ---------------------------------------------------------------------------
# TTN
# ---------------------------------------------------------------------------


class _Tree:
    """Rooted view of a loop-free graph: every non-root node has two children
    (nodes or sites, ordered by their first site) in slots 0, 1 and its parent
    in slot 2; the root has three children."""

    def __init__(self, N: int, graph, sitenodes, root: Node):
        self.N, self.graph, self.sitenodes, self.root = N, graph, dict(sitenodes), root
        self.nodes = sorted(graph)
        self.index = {n: k for k, n in enumerate(self.nodes)}
        _, par = _bfs(graph, root)
        self.parent = {n: par.get(n) for n in graph}
        kids = {n: [] for n in graph}
        for i in range(1, N + 1):
            kids[sitenodes[i]].append(("site", i))
        for n in graph:
            if self.parent[n] is not None:
                kids[self.parent[n]].append(("node", n))
        self.below: Dict[tuple, List[int]] = {}

        def below(it):
            if it not in self.below:
                self.below[it] = [it[1]] if it[0] == "site" else sorted(x for k in kids[it[1]] for x in below(k))
            return self.below[it]

        self.slots: Dict[Node, List[tuple]] = {}
        for n in graph:
            ks = sorted(kids[n], key=lambda k: below(k)[0])
            want = 3 if self.parent[n] is None else 2
            if len(ks) != want:
                raise ValueError(f"TTN: node {n} must have three legs")
            self.slots[n] = ks if self.parent[n] is None else ks + [("node", self.parent[n])]
        for n in graph:
            below(("node", n))

    def slot_of(self, n: Node, m: Node) -> int:
        return self.slots[n].index(("node", m))

    def chan_item(self, n: Node, s: int) -> tuple:
        """Subtree whose channels label the link at slot s of n."""
        return self.slots[n][s] if (s < 2 or self.parent[n] is None) else ("node", n)


class TTN:
    """TTN{T}(sites, graph, tensors, orthocenter) (src/ttn/ttn.jl:1-30), dense
    FP64 node tensors on the host, legs in slot order."""

    def __init__(self, tree: _Tree, tensors: Dict[Node, np.ndarray], orthocenter: Optional[Node]):
        self.tree = tree
        self.tensors = tensors
        self.orthocenter = orthocenter

    @property
    def graph(self):
        return self.tree.graph

    def numsites(self) -> int:
        return self.tree.N

    def __getitem__(self, node: Node) -> np.ndarray:
        return self.tensors[node]

    def copy(self) -> "TTN":
        return TTN(self.tree, {k: v.copy() for k, v in self.tensors.items()}, self.orthocenter)

    def maxlinkdim(self) -> int:
        """maxlinkdim(ttn) (ttn.jl:180-195)."""
        md = 1
        for n, A in self.tensors.items():
            for s, it in enumerate(self.tree.slots[n]):
                if it[0] == "node":
                    md = max(md, A.shape[s])
        return md


maxlinkdim = TTN.maxlinkdim


def _structure_handle(ctx: Context, tree: _Tree, nchan=None):
    h = C.c_void_p()
    ctx.check(ctx.lib.tn_ttn_create(ctx.h, len(tree.nodes), C.byref(h)))
    for n in tree.nodes:
        for s, it in enumerate(tree.slots[n]):
            if it[0] == "node" and tree.index[n] < tree.index[it[1]]:
                m = it[1]
                k = 1 if nchan is None else nchan[tree.chan_item(n, s)]
                ctx.check(ctx.lib.tn_ttn_set_link(h, tree.index[n], s, tree.index[m], tree.slot_of(m, n), k))
    return h


def _upload(ctx, h, tree: _Tree, tensors):
    for n in tree.nodes:
        A = np.ascontiguousarray(tensors[n], dtype=np.float64)
        dims = (C.c_long * 3)(*A.shape)
        ctx.check(ctx.lib.tn_ttn_set_node(h, tree.index[n], dims, A.ctypes.data_as(L._PD)))


def _download(ctx, h, tree: _Tree) -> Dict[Node, np.ndarray]:
    out = {}
    for n in tree.nodes:
        dims = (C.c_long * 3)()
        ctx.check(ctx.lib.tn_ttn_get_node(h, tree.index[n], dims, None))
        A = np.empty(tuple(dims), dtype=np.float64)
        ctx.check(ctx.lib.tn_ttn_get_node(h, tree.index[n], dims, A.ctypes.data_as(L._PD)))
        out[n] = A
    return out


def _move(ctx, h, tree, a: Node, b: Node, maxdim=TYPEMAX_INT, mindim=1, cutoff=-1.0):
    """moveisometry_to_next!(ttn, a, b; ...) on the device (ttn.jl:266-310)."""
    terr, nd = C.c_double(), C.c_long()
    use_cut = cutoff if cutoff > 0.0 else -1.0
    ctx.check(ctx.lib.tn_ttn_move(h, tree.index[a], tree.slot_of(a, b), int(min(maxdim, TYPEMAX_INT)), int(mindim),
                                  float(use_cut), C.byref(terr), C.byref(nd)))
    return terr.value


def randomTTN(numsites: int, graph, sitenodes, chi: int, seed: int = 0, ctx: Optional[Context] = None) -> TTN:
    """randomTTN(sites, graph, sitenodes, chi) (ttn_generators.jl:200-262) for
    spin-1/2 sites without QNs: link dims min(product of the node's other
    legs, chi), Gaussian entries (numpy default_rng(seed), nodes in sorted
    order), isometrised on the GPU toward the eccentric center."""
    for b in range(1, numsites + 1):
        if b not in sitenodes:
            raise ValueError(f'`randomTTN()`: `sitenodes` does not have site "{b}" !!')
    center = find_eccentric_central_node(graph, set(sitenodes.values()))
    tree = _Tree(numsites, graph, sitenodes, center)
    dims = {("site", i): 2 for i in range(1, numsites + 1)}
    dist, _ = _bfs(graph, center)
    for n in sorted(graph, key=lambda x: (-dist[x], x)):
        if n != center:
            a, b = tree.slots[n][:2]
            dims[("node", n)] = min(dims[a] * dims[b], chi)
    rng = np.random.default_rng(seed)
    T = {}
    for n in tree.nodes:
        A = rng.standard_normal(tuple(dims[tree.chan_item(n, s)] for s in range(3)))
        T[n] = A / np.linalg.norm(A)
    ctx = ctx or default_context()
    h = _structure_handle(ctx, tree)
    try:
        _upload(ctx, h, tree, T)
        for n in sorted(graph, key=lambda x: (-dist[x], x)):
            if n != center:
                _move(ctx, h, tree, n, nextnode_in_path(graph, n, center), maxdim=chi, cutoff=0.0)
        ctx.check(ctx.lib.tn_ttn_normalize(h, tree.index[center], None))
        T = _download(ctx, h, tree)
    finally:
        ctx.lib.tn_ttn_destroy(h)
    return TTN(tree, T, center)


def default_randomTTN(numsites: int, chi: int, seed: int = 0, ctx: Optional[Context] = None) -> TTN:
    """default_randomTTN(sites, chi) (ttn_generators.jl:264-279)."""
    graph, sitenodes = default_graph_sitenodes(numsites)
    return randomTTN(numsites, graph, sitenodes, chi, seed, ctx)


def default_sweeppath(psi: TTN) -> List[Node]:
    """default_sweeppath(psi) (sweep_ttn.jl:40-60)."""
    nsites = _pow2_ge(psi.numsites())
    nlayers = _tz(nsites)
    path = []
    for ll in range(nlayers - 1, 0, -1):
        cnt = nsites >> ll
        for nn in range(1, cnt + 1):
            node = (ll, nn if (nlayers - ll) % 2 == 1 else cnt - nn + 1)
            if ll == 1 and node not in psi.graph:
                continue
            path.append(node)
    return path


def _tree_operator(tree: _Tree, H: CouplingModel):
    """Channels of every link and the sparse coupling of every node."""
    onebody, twobody = H._few_body()
    partners: Dict[Tuple[int, str], List[int]] = {}
    for (i, o1, j, o2) in twobody:
        partners.setdefault((i, o1), []).append(j)
        partners.setdefault((j, o2), []).append(i)
    chan: Dict[tuple, List[tuple]] = {}
    for it, sites in tree.below.items():
        S = set(sites)
        ch = [("I",), ("H",)]
        for i in sites:
            for o in sorted(o for (j, o) in partners if j == i):
                if any(k not in S for k in partners[(i, o)]):
                    ch.append(("op", i, o))
        chan[it] = ch
    W: Dict[Node, List[tuple]] = {}
    for n in tree.nodes:
        items = [tree.chan_item(n, s) for s in range(3)]
        idx = [{c: k for k, c in enumerate(chan[it])} for it in items]
        S = [set(tree.below[it]) for it in items]
        I = [ix[("I",)] for ix in idx]
        Hh = [ix[("H",)] for ix in idx]
        acc: Dict[Tuple[int, int, int], float] = {}

        def add(c, w):
            acc[tuple(c)] = acc.get(tuple(c), 0.0) + w

        root = tree.parent[n] is None
        if root:
            add((Hh[0], I[1], I[2]), 1.0)
            add((I[0], Hh[1], I[2]), 1.0)
            add((I[0], I[1], Hh[2]), 1.0)
        else:
            add((I[0], I[1], I[2]), 1.0)
            add((Hh[0], I[1], Hh[2]), 1.0)
            add((I[0], Hh[1], Hh[2]), 1.0)
        legs = 3 if root else 2
        for (i, o1, j, o2), J in twobody.items():
            a = [k for k in range(legs) if i in S[k]]
            b = [k for k in range(legs) if j in S[k]]
            if a and b and a[0] != b[0]:
                c = list(I) if root else [I[0], I[1], Hh[2]]
                c[a[0]] = idx[a[0]][("op", i, o1)]
                c[b[0]] = idx[b[0]][("op", j, o2)]
                add(c, J)
        if not root:
            for k in range(2):
                for ch in chan[items[k]]:
                    if ch[0] == "op" and ch in idx[2]:
                        c = [I[0], I[1], idx[2][ch]]
                        c[k] = idx[k][ch]
                        add(c, 1.0)
        W[n] = sorted((c[0], c[1], c[2], w) for c, w in acc.items())
    site_ops = {}
    for i in range(1, tree.N + 1):
        ch = chan[("site", i)]
        F = np.zeros((len(ch), 2, 2))
        for k, c in enumerate(ch):
            F[k] = np.eye(2) if c[0] == "I" else onebody[i] if c[0] == "H" else op_matrix(c[2])
        site_ops[i] = F  # [c][bra][ket]
    return chan, W, site_ops


# ---------------------------------------------------------------------------
# StateEnvsTTN
# ---------------------------------------------------------------------------


class StateEnvsTTN:
    """StateEnvsTTN(psi, H::CouplingModel) (state_envs_ttn.jl:20-30): copies
    psi to the GPU and builds the environments toward its orthogonality
    center (LinkTensorsTTN(psi, M), linktensors.jl:240-262).
    StateEnvsTTN(psi, H, Ms; weight) (state_envs_ttn.jl:36-70) adds the
    penalty projectors of the TTNs Ms (EnvCouplingModelProjTTN,
    environment.jl:42-110), for excited states."""

    def __init__(self, psi: TTN, H: CouplingModel, Ms: Optional[Sequence[TTN]] = None,
                 weight: Optional[float] = None, ctx: Optional[Context] = None):
        if Ms is not None:
            if weight is None:
                raise TypeError("StateEnvsTTN(psi, H, Ms; weight): keyword `weight` is required")
            if weight <= 0.0:
                raise ValueError(f"`weight` parameter should be > 0.0 (value passed was `weight={weight}`)")
            for M in Ms:
                if (M.tree.graph != psi.tree.graph or M.tree.sitenodes != psi.tree.sitenodes
                        or M.tree.root != psi.tree.root):
                    raise ValueError("StateEnvsTTN: the projector TTNs must share psi's graph")
        if H.N != psi.numsites():
            raise ValueError("StateEnvsTTN: the CouplingModel and the TTN have different sites")
        if psi.orthocenter is None:
            raise ValueError("`LinkTensorsTTN()`: TTN does not have a proper orthogonality center !!")
        self.ctx = ctx or default_context()
        self.tree = psi.tree
        self.H = H
        chan, W, site_ops = _tree_operator(self.tree, H)
        nchan = {it: len(c) for it, c in chan.items()}
        self.h = _structure_handle(self.ctx, self.tree, nchan)
        _live_children.add(self)
        lib, t = self.ctx.lib, self.tree
        for n in t.nodes:
            for s, it in enumerate(t.slots[n]):
                if it[0] == "site":
                    F = np.ascontiguousarray(site_ops[it[1]])
                    self.ctx.check(lib.tn_ttn_set_site(self.h, t.index[n], s, 2, F.shape[0], F.ctypes.data_as(L._PD)))
        _upload(self.ctx, self.h, t, psi.tensors)
        if Ms is not None:
            self.ctx.check(lib.tn_ttn_set_projectors(self.h, len(Ms), float(weight)))
            for j, M in enumerate(Ms):
                for n in t.nodes:
                    A = np.ascontiguousarray(M.tensors[n], dtype=np.float64)
                    dims = (C.c_long * 3)(*A.shape)
                    self.ctx.check(lib.tn_ttn_set_projector_node(self.h, j, t.index[n], dims,
                                                                 A.ctypes.data_as(L._PD)))
        for n in t.nodes:
            w = W[n]
            c0 = (C.c_int * len(w))(*[x[0] for x in w])
            c1 = (C.c_int * len(w))(*[x[1] for x in w])
            c2 = (C.c_int * len(w))(*[x[2] for x in w])
            ww = (C.c_double * len(w))(*[x[3] for x in w])
            self.ctx.check(lib.tn_ttn_set_coupling(self.h, t.index[n], len(w), c0, c1, c2, ww))
        dist, _ = _bfs(t.graph, psi.orthocenter)
        for n in sorted(t.graph, key=lambda x: (-dist[x], x)):
            if n != psi.orthocenter:
                self._env(n, nextnode_in_path(t.graph, n, psi.orthocenter))
        self.ctx.check(lib.tn_ttn_set_center(self.h, t.index[psi.orthocenter]))
        self.ctx.check(lib.tn_ttn_seed(self.h, 0))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.lib.tn_ttn_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _env(self, a: Node, b: Node):
        self.ctx.check(self.ctx.lib.tn_ttn_update_env(self.h, self.tree.index[a], self.tree.slot_of(a, b)))

    @property
    def orthocenter(self) -> Node:
        k = C.c_int()
        self.ctx.check(self.ctx.lib.tn_ttn_get_center(self.h, C.byref(k)))
        return self.tree.nodes[k.value] if k.value >= 0 else None

    @orthocenter.setter
    def orthocenter(self, node: Node):
        self.ctx.check(self.ctx.lib.tn_ttn_set_center(self.h, self.tree.index[node]))

    @property
    def psi(self) -> TTN:
        return TTN(self.tree, _download(self.ctx, self.h, self.tree), self.orthocenter)

    def node_shape(self, node: Node) -> Tuple[int, int, int]:
        dims = (C.c_long * 3)()
        self.ctx.check(self.ctx.lib.tn_ttn_get_node(self.h, self.tree.index[node], dims, None))
        return tuple(dims)


def getpsi(sysenv: StateEnvsTTN) -> TTN:
    return sysenv.psi


def position_(sysenv: StateEnvsTTN, node: Node, *, maxdim: int = TYPEMAX_INT, mindim: int = 1,
              cutoff: float = FLOAT64_THRESHOLD, svd_alg: str = "divide_and_conquer", normalize: bool = True,
              node_to_skip: Optional[Node] = None) -> None:
    """position!(sysenv, node; ...) (state_envs_ttn.jl:120-140): isometrize!
    (ttn.jl:342-382) then move_environment! along the same path."""
    t = sysenv.tree
    if node not in t.graph:
        raise ValueError("`isometrize!()`: Input position is ill-defined !!")
    skip = t.index[node_to_skip] if node_to_skip is not None else -1
    sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_position(sysenv.h, t.index[node], int(min(maxdim, TYPEMAX_INT)),
                                                    int(mindim), float(cutoff), int(normalize), skip))


def product(sysenv: StateEnvsTTN, v: np.ndarray) -> np.ndarray:
    """product(sysenv, v) at the orthogonality center (state_envs_ttn.jl:150-155)."""
    from .core import DeviceArray
    node = sysenv.orthocenter
    shape = sysenv.node_shape(node)
    v = np.ascontiguousarray(v, dtype=np.float64)
    if v.size != int(np.prod(shape)):
        raise ValueError("The order of the LinkTensorsTTN-ITensor product P*v is not equal to the order of the "
                         "ITensor v, this is probably due to an index mismatch.")
    dv = DeviceArray.from_numpy(sysenv.ctx, v.reshape(shape))
    do = DeviceArray(sysenv.ctx, shape)
    sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_matvec(sysenv.h, sysenv.tree.index[node], dv.ptr, do.ptr))
    return do.numpy()


def update_position_(sysenv: StateEnvsTTN, solver, node: Node, *, time_step=None, normalize: bool = True,
                     maxdim: int = TYPEMAX_INT, mindim: int = 1, cutoff: float = FLOAT64_THRESHOLD,
                     svd_alg: str = "divide_and_conquer", **kwargs) -> float:
    """update_position!(sysenv, solver, node; ...) (update_site_ttn.jl:42-63)."""
    if solver is not eig_solver:
        raise NotImplementedError("TTN updates support `eig_solver` only (as the reference)")
    if node not in sysenv.tree.graph:
        raise ValueError("`isometrize!()`: Input position is ill-defined !!")
    e = C.c_double()
    sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_update_position(
        sysenv.h, sysenv.tree.index[node], int(min(maxdim, TYPEMAX_INT)), int(mindim), float(cutoff),
        int(normalize), int(kwargs.get("solver_krylovdim", 5)), int(kwargs.get("solver_maxiter", 2)),
        float(kwargs.get("solver_tol", 1e-14)), int(bool(kwargs.get("solver_check_convergence", False))),
        C.byref(e)))
    return e.value


def subspace_expand_(sysenv: StateEnvsTTN, node: Node, nextnode: Node, max_expand_dim: int,
                     noise: float) -> None:
    """subspace_expand!(psi, node, nextnode, max_expand_dim, noise)
    (update_site_ttn.jl:65-109), dense: both tensors grow along their common
    bond by p = min(sum of each tensor's other leg dims, max_expand_dim),
    padded with Gaussian entries (the StateEnvsTTN's seeded generator)
    scaled to noise * |tensor|, nextnode's padding first."""
    t = sysenv.tree
    sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_expand_random(sysenv.h, t.index[node], t.slot_of(node, nextnode),
                                                         int(max_expand_dim), float(noise)))


@dataclass
class SweepDataTTN:
    """SweepDataTTN (sweep_ttn.jl:1-30)."""
    sweepcount: int = 0
    maxchi: List[int] = field(default_factory=list)
    energy: List[float] = field(default_factory=list)


def _maxlinkdim(sysenv: StateEnvsTTN) -> int:
    md = 1
    for n in sysenv.tree.nodes:
        dims = sysenv.node_shape(n)
        for s, it in enumerate(sysenv.tree.slots[n]):
            if it[0] == "node":
                md = max(md, dims[s])
    return md


def fullsweep_(sysenv: StateEnvsTTN, sweeppath: Sequence[Node], solver, swdata: SweepDataTTN, **kwargs) -> float:
    """fullsweep!(sysenv, sweeppath, solver, swdata; kwargs...) (sweep_ttn.jl:117-274)."""
    if set(sweeppath) != set(sysenv.tree.graph):
        raise ValueError("fullsweep!: the sweep path must contain every node")
    outputlevel = kwargs.get("outputlevel", 1)
    maxdim = int(kwargs.get("maxdim", TYPEMAX_INT))
    mindim = int(kwargs.get("mindim", 1))
    cutoff = float(kwargs.get("cutoff", FLOAT64_THRESHOLD))
    normalize = bool(kwargs.get("normalize", True))
    noise = float(kwargs.get("noise", 0.0))
    expand_dim = int(kwargs.get("expand_dim", 0 if abs(noise) < 100 * FLOAT64_THRESHOLD else 20))
    max_expand_dim = int(kwargs.get("max_expand_dim", 2 * expand_dim))
    expand_numiter = int(kwargs.get("expand_numiter", 4))
    if expand_dim != 0 and expand_numiter < 2:
        raise ValueError(f"`fullsweep!()`: `expand_numiter={expand_numiter}` cannot be less than 2 "
                         f"for `expand_dim={expand_dim}` !!")
    skw = {k: kwargs[k] for k in ("solver_krylovdim", "solver_maxiter", "solver_tol", "solver_check_convergence")
           if k in kwargs}
    energy = float("nan")
    swdata.sweepcount += 1
    path = list(sweeppath)
    t = sysenv.tree
    if outputlevel <= 1:  # the whole sweep in the library
        central = find_eccentric_central_node(t.graph)
        ids = (C.c_int * len(path))(*[t.index[n] for n in path])
        e, mc = C.c_double(), C.c_long()
        sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_fullsweep(
            sysenv.h, ids, len(path), t.index[central], int(min(maxdim, TYPEMAX_INT)), mindim, cutoff, noise,
            expand_dim, max_expand_dim, expand_numiter, int(skw.get("solver_krylovdim", 5)),
            int(skw.get("solver_maxiter", 2)), float(skw.get("solver_tol", 1e-14)), int(normalize),
            int(bool(skw.get("solver_check_convergence", False))), C.byref(e), C.byref(mc)))
        energy = e.value
        swdata.maxchi.append(int(mc.value))
    elif abs(noise) < 100 * FLOAT64_THRESHOLD:
        for node in path + path[::-1]:
            energy = update_position_(sysenv, solver, node, normalize=normalize, maxdim=maxdim, mindim=mindim,
                                      cutoff=-1.0, **skw)
            if outputlevel > 1:
                print(f"At sweep {swdata.sweepcount} updated node {node} => Energy {energy}")
    else:
        central = find_eccentric_central_node(sysenv.tree.graph)
        for ii in list(range(len(path))) + list(range(len(path) - 1, -1, -1)):
            node = path[ii]
            if node == central:
                energy = update_position_(sysenv, solver, node, normalize=normalize, maxdim=maxdim, mindim=mindim,
                                          cutoff=-1.0, **skw)
                continue
            nxt = nextnode_in_path(sysenv.tree.graph, node, central)
            position_(sysenv, node, normalize=normalize, maxdim=maxdim, mindim=mindim, cutoff=-1.0,
                      node_to_skip=nxt)
            subspace_expand_(sysenv, node, nxt, max_expand_dim, noise)
            sysenv.orthocenter = nxt
            for dummy in range(1, expand_numiter + 1):
                md = maxdim if dummy == expand_numiter else (maxdim + max_expand_dim if dummy == 1
                                                             else maxdim + expand_dim)
                newnode = node if dummy % 2 == 1 else nxt
                energy = update_position_(sysenv, solver, newnode, normalize=normalize, maxdim=md, mindim=mindim,
                                          cutoff=cutoff, **skw)
    if outputlevel > 1:
        swdata.maxchi.append(_maxlinkdim(sysenv))
    swdata.energy.append(energy)
    enerr = swdata.energy[-1] - swdata.energy[-2] if swdata.sweepcount > 1 else float("nan")
    if outputlevel > 0:
        print(f"At sweep {swdata.sweepcount} => E={energy}, MaxLinkDim={swdata.maxchi[-1]}, "
              f"Noise={noise:.2g}")
    return enerr


@dataclass
class OptimizeParamsTTN:
    """OptimizeParamsTTN (optimize_ttn.jl:1-100)."""
    maxdim: List[int]
    nsweeps: List[int]
    cutoff: List[float]
    noise: List[float]
    noisedecay: List[float]
    disable_noise_after: List[int]


def make_optimize_params(*, maxdim, nsweeps, cutoff=FLOAT64_THRESHOLD, noise=0.0, noisedecay=1.0,
                         disable_noise_after=TYPEMAX_INT) -> OptimizeParamsTTN:
    """OptimizeParamsTTN(; maxdim, nsweeps, cutoff, noise, noisedecay,
    disable_noise_after) (optimize_ttn.jl:60-100)."""
    n = len(nsweeps)
    vec = lambda x, cast: [cast(v) for v in x] if isinstance(x, (list, tuple)) else [cast(x)] * n
    p = OptimizeParamsTTN(list(maxdim), list(nsweeps), vec(cutoff, float), vec(noise, float),
                          vec(noisedecay, float), vec(disable_noise_after, int))
    if not (len(p.nsweeps) == len(p.maxdim) == len(p.cutoff) == len(p.noise) == len(p.noisedecay)
            == len(p.disable_noise_after)):
        raise ValueError("`OptimizeParamsTTN()` :: Size mismatch in input vectors !! \n"
                         "Lengths of `maxdim` and `nsweeps` must be same !!\n"
                         "`cutoff`, `noise`, `noisedecay`, `disable_noise_after` "
                         "can be either scalar or vector having length = `length(maxdim)` !!")
    return p


def optimize_(sysenv: StateEnvsTTN, params: OptimizeParamsTTN, sweeppath: Sequence[Node], **kwargs) -> SweepDataTTN:
    """optimize!(sysenv, params, sweeppath; kwargs...) (optimize_ttn.jl:148-223)."""
    enerrgoal = kwargs.pop("energyErrGoal", None)
    seed = kwargs.pop("seed", None)
    if seed is not None:
        sysenv.ctx.check(sysenv.ctx.lib.tn_ttn_seed(sysenv.h, int(seed)))
    sw = SweepDataTTN()
    for ii in range(len(params.nsweeps)):
        noise = params.noise[ii]
        for jj in range(1, params.nsweeps[ii] + 1):
            enerr = fullsweep_(sysenv, sweeppath, eig_solver, sw, maxdim=params.maxdim[ii], cutoff=params.cutoff[ii],
                               noise=noise, **kwargs)
            if enerrgoal is not None and abs(enerr) < abs(enerrgoal) and abs(noise) < FLOAT64_THRESHOLD:
                break
            if jj == params.disable_noise_after[ii]:
                noise = 0.0
            noise /= params.noisedecay[ii]
            if noise < 100 * FLOAT64_THRESHOLD:
                noise = 0.0
    return sw


def optimize(psi0: TTN, H: CouplingModel, *args, **kwargs):
    """optimize(psi0, H, params, sweeppath; kwargs...) (optimize_ttn.jl:225-260)
    or, for excited states, optimize(psi0, H, Ms, params, sweeppath; weight,
    kwargs...) (optimize_ttn.jl:284-300) -> (energy, psi)."""
    if args and isinstance(args[0], OptimizeParamsTTN):
        Ms, (params, sweeppath) = None, args[:2]
        sysenv = StateEnvsTTN(psi0, H)
    else:
        Ms, params, sweeppath = args[:3]
        sysenv = StateEnvsTTN(psi0, H, Ms, weight=kwargs.pop("weight", -1.0))
    try:
        sw = optimize_(sysenv, params, sweeppath, **kwargs)
        return sw.energy[-1], sysenv.psi
    finally:
        sysenv.close()
