"""Pins of the TTN oracle (oracle/tn_oracle_ttn.py) -- CPU only.

The tree operator and the environment moves are checked against the dense
Hamiltonian (Kronecker construction) projected on the tree: for a random
isometric TTN and every node as center, <psi(x)|H|psi(y)> equals
<x, H_eff y>, where psi(x) is the dense state with the center tensor
replaced by x.  Ground-state optimisation is pinned against exact
diagonalisation.  Graph helpers follow the examples in the reference's own
docstrings.
"""

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from oracle import tn_oracle_ttn as TO

E0_N12 = -5.14209063284054  # Heisenberg OBC N = 12 (ED)


def dense_h(terms: TO.CouplingTerms):
    N = terms.N
    H = sps.csr_matrix((2 ** N, 2 ** N))
    for coef, ops in terms.terms:
        mats = [sps.identity(2, format="csr")] * N
        mats = list(mats)
        for o, i in ops:
            mats[i - 1] = sps.csr_matrix(TO.SPIN_HALF[o]) @ mats[i - 1]
        M = mats[0]
        for x in mats[1:]:
            M = sps.kron(M, x, format="csr")
        H = H + coef * M
    return H


def ttn_to_dense(psi: TO.TTNState, center=None, x=None):
    """Dense state (sites in chain order) of a TTN; optionally with the
    tensor at `center` replaced by x."""
    t = psi.tree
    T = dict(psi.T)
    if center is not None:
        T[center] = x

    def sub(n, up):
        """(2^k x D_up) map of the subtree of n away from `up`, and its sites."""
        A = T[n]
        sites, up_slot = [], None
        for s, it in enumerate(t.slots[n]):
            if it[0] == "node" and it[1] == up:
                up_slot = s
                continue
            if it[0] == "site":
                sites.append(it[1])
                continue
            V, st = sub(it[1], n)
            A = np.moveaxis(np.tensordot(V, A, axes=([1], [s])), 0, s)
            sites += st
        if up_slot is None:
            return A.reshape(-1, 1), sites
        A = np.moveaxis(A, up_slot, -1)
        return A.reshape(-1, A.shape[-1]), sites

    v, sites = sub(t.root, None)
    v = v.reshape([2] * len(sites))
    return np.transpose(v, np.argsort(sites)).reshape(-1)


@pytest.mark.parametrize("N", [5, 6, 7, 12, 13, 16, 23, 32, 48, 64])
def test_default_graph_structure(N):
    edges, sitenodes = TO.default_graph_sitenodes(N)
    assert sorted(sitenodes) == list(range(1, N + 1))
    t = TO.Tree.build(N)
    for n in edges:
        assert len(t.slots[n]) == 3
    path = TO.default_sweeppath(N, edges)
    assert set(path) == set(edges) and len(path) == len(edges)


def test_default_sweeppath_hand_enumerated():
    """Paths worked out by hand from sweep_ttn.jl:49-65 and ttn_generators.jl:3-94.

    N = 8: 8 sites in pairs on (1,1)..(1,4); nlayers = 3.  Layer 2 runs
    forward ((3 - 2) odd), layer 1 backward ((3 - 1) even).
    N = 13: _distribute_site_positions gives the pairs (1,0) (1,1) (1,0) (1,1)
    (1,1) (1,1) (1,0) (1,1); a pair with a 0 has no layer-1 node (its site
    hangs on layer 2), so layer 1 holds (1,2) (1,4) (1,5) (1,6) (1,8).  With
    nlayers = 4, layer 3 runs forward, layer 2 backward, layer 1 forward with
    the missing nodes skipped."""
    edges8, sn8 = TO.default_graph_sitenodes(8)
    assert TO.default_sweeppath(8, edges8) == [(2, 1), (2, 2), (1, 4), (1, 3), (1, 2), (1, 1)]
    edges13, sn13 = TO.default_graph_sitenodes(13)
    assert TO.default_sweeppath(13, edges13) == [(3, 1), (3, 2), (2, 4), (2, 3), (2, 2), (2, 1),
                                                 (1, 2), (1, 4), (1, 5), (1, 6), (1, 8)]
    # sites on layer 2 where a layer-1 pair is half empty
    assert sn13[1] == (2, 1) and sn13[2] == (1, 2) and sn13[3] == (1, 2) and sn13[4] == (2, 2)
    assert sn13[10] == (1, 6) and sn13[11] == (2, 4) and sn13[12] == (1, 8) and sn13[13] == (1, 8)


def test_default_graph_docstring_examples():
    # ttn_generators.jl:49-57
    _, sn = TO.default_graph_sitenodes(32)
    assert sn[1] == (1, 1) and sn[2] == (1, 1) and sn[3] == (1, 2) and sn[4] == (1, 2)
    assert sn[31] == (1, 16) and sn[32] == (1, 16)


def test_minimum_power2():
    assert [TO.min_pow2_ge(n) for n in (1, 2, 3, 5, 16, 17, 33)] == [1, 2, 4, 8, 16, 32, 64]


@pytest.mark.parametrize("N,chi", [(7, 4), (10, 8), (12, 6)])
def test_effective_hamiltonian_is_projected_dense_hamiltonian(N, chi):
    rng = np.random.default_rng(N)
    H = TO.heisenberg_terms(N, 1.0, rng.uniform(-2, 2, N))
    Hd = dense_h(H)
    t = TO.Tree.build(N)
    psi = TO.random_ttn(t, chi, seed=N)
    env = TO.StateEnvsTTN(psi, H)
    for node in TO.default_sweeppath(N, t.edges):
        env.position(node)
        A = env.psi.T[node]
        x = rng.standard_normal(A.shape)
        y = rng.standard_normal(A.shape)
        px = ttn_to_dense(env.psi, node, x)
        py = ttn_to_dense(env.psi, node, y)
        assert abs(px @ py - np.vdot(x, y)) < 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)  # isometric
        want = px @ (Hd @ py)
        got = np.vdot(x, env.env.product(node, y))
        assert abs(got - want) < 1e-11 * max(1.0, abs(want)), node


def test_isometry_moves_keep_the_state():
    N = 10
    t = TO.Tree.build(N)
    psi = TO.random_ttn(t, 8, seed=3)
    ref = ttn_to_dense(psi)
    assert abs(np.linalg.norm(ref) - 1) < 1e-12
    for node in TO.default_sweeppath(N, t.edges):
        TO.isometrize(psi, node)
        assert np.abs(ttn_to_dense(psi) - ref).max() < 1e-12


def test_optimize_reaches_ed_n12():
    N = 12
    H = TO.heisenberg_terms(N)
    t = TO.Tree.build(N)
    env = TO.StateEnvsTTN(TO.random_ttn(t, 64, seed=1), H)
    path = TO.default_sweeppath(N, t.edges)
    sw = TO.optimize(env, TO.OptimizeParamsTTN.make([64], [4]), path)
    assert abs(sw.energy[-1] - E0_N12) < 1e-10
    psi = ttn_to_dense(env.psi)
    assert abs(psi @ (dense_h(H) @ psi) - E0_N12) < 1e-10


def test_noisy_optimize_reference_settings_reaches_ed():
    """test/test_TTN.jl:27-38: default_randomTTN(sites, 12), maxdim 128,
    5 sweeps, cutoff 1e-14, noise 1e-2, noisedecay 5, disable_noise_after 2."""
    N = 12
    H = TO.heisenberg_terms(N)
    t = TO.Tree.build(N)
    env = TO.StateEnvsTTN(TO.random_ttn(t, 12, seed=2), H)
    path = TO.default_sweeppath(N, t.edges)
    p = TO.OptimizeParamsTTN.make([128], [5], cutoff=1e-14, noise=1e-2, noisedecay=5.0, disable_noise_after=2)
    sw = TO.optimize(env, p, path, seed=5)
    assert abs(sw.energy[-1] - E0_N12) < 1e-9
    assert sw.maxchi[-1] > 12  # the bonds grew through subspace expansion


def test_random_field_instance_matches_ed():
    N = 14
    rng = np.random.default_rng(5)
    H = TO.heisenberg_terms(N, 1.0, rng.uniform(-2, 2, N))
    e0 = spla.eigsh(dense_h(H), k=1, which="SA", tol=1e-12)[0][0]
    t = TO.Tree.build(N)
    env = TO.StateEnvsTTN(TO.random_ttn(t, 128, seed=4), H)
    sw = TO.optimize(env, TO.OptimizeParamsTTN.make([128], [4]), TO.default_sweeppath(N, t.edges))
    assert abs(sw.energy[-1] - e0) < 1e-9


@pytest.mark.parametrize("N,chi", [(7, 4), (12, 6)])
def test_projector_operator_is_projected_dense_operator(N, chi):
    """H + w sum_k |M_k><M_k| (LinkProjTTN, linkproj.jl) projected on the
    tree at every node equals the node product with projectors; the M_k are
    random TTNs of other bond dimensions on the same tree."""
    rng = np.random.default_rng(100 + N)
    H = TO.heisenberg_terms(N, 1.0, rng.uniform(-1, 1, N))
    t = TO.Tree.build(N)
    Ms = [TO.random_ttn(t, chi + 1 + k, seed=200 + k) for k in range(2)]
    w = 3.5
    Hd = dense_h(H).toarray()
    for M in Ms:
        m = ttn_to_dense(M)
        Hd = Hd + w * np.outer(m, m) / (m @ m)
    Ms = [M.copy() for M in Ms]
    for M in Ms:  # unit norm, as optimised states are
        nrm = np.sqrt(ttn_to_dense(M) @ ttn_to_dense(M))
        M.T[M.center] = M.T[M.center] / nrm
    env = TO.StateEnvsTTN(TO.random_ttn(t, chi, seed=N), H, Ms, weight=w)
    for node in TO.default_sweeppath(N, t.edges):
        env.position(node)
        A = env.psi.T[node]
        x = rng.standard_normal(A.shape)
        y = rng.standard_normal(A.shape)
        want = ttn_to_dense(env.psi, node, x) @ (Hd @ ttn_to_dense(env.psi, node, y))
        got = np.vdot(x, env.env.product(node, y))
        assert abs(got - want) < 1e-11 * max(1.0, abs(want)), node
    with pytest.raises(ValueError, match="should be > 0.0"):
        TO.StateEnvsTTN(env.psi, H, Ms, weight=0.0)


def test_excited_state_optimize_reference_settings_reaches_first_excited_level():
    """test_TTN.jl:40-64 without quantum numbers: the ground state from the
    reference settings, then optimize(psi0, H, [psi_gr], params; weight = 10)
    with maxdim 128, 10 sweeps, noise 1e-2 (decay 5, off after 6): the lowest
    triplet, E1 = -4.86114793703639 (ED), orthogonal to the ground state."""
    N = 12
    H = TO.heisenberg_terms(N, 1.0)
    t = TO.Tree.build(N)
    path = TO.default_sweeppath(N, t.edges)
    gs = TO.StateEnvsTTN(TO.random_ttn(t, 12, seed=1), H)
    TO.optimize(gs, TO.OptimizeParamsTTN.make([128], [5], cutoff=1e-14, noise=1e-2, noisedecay=5,
                                              disable_noise_after=2), path)
    ex = TO.StateEnvsTTN(TO.random_ttn(t, 12, seed=2), H, [gs.psi], weight=10.0)
    sw = TO.optimize(ex, TO.OptimizeParamsTTN.make([128], [10], cutoff=1e-14, noise=1e-2, noisedecay=5,
                                                   disable_noise_after=6), path)
    assert abs(sw.energy[-1] - (-4.86114793703639)) < 1e-8
    assert abs(ttn_to_dense(gs.psi) @ ttn_to_dense(ex.psi)) < 1e-5
