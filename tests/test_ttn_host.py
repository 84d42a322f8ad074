"""Host logic of the TTN path (graph, sweep path, tree operator) against the
oracle -- CPU only, no GPU calls."""

import numpy as np
import pytest

from oracle import tn_oracle_ttn as TO
import tennetlib_amd as T
from tennetlib_amd import ttn as TT


@pytest.mark.parametrize("N", [5, 7, 12, 16, 23, 40, 64])
def test_graph_and_sweeppath_match_oracle(N):
    g, sn = T.default_graph_sitenodes(N)
    og, osn = TO.default_graph_sitenodes(N)
    assert g == og and sn == osn
    center = T.find_eccentric_central_node(g, set(sn.values()))
    assert center == TO.eccentric_center(og, set(osn.values()))
    tree = TT._Tree(N, g, sn, center)
    otree = TO.Tree.build(N)
    assert tree.slots == otree.slots
    psi = T.TTN(tree, {}, center)
    assert T.default_sweeppath(psi) == TO.default_sweeppath(N, og)


@pytest.mark.parametrize("N", [7, 12, 13])
def test_tree_operator_matches_oracle(N):
    h = np.random.default_rng(N).uniform(-2, 2, N)
    os = T.OpStrings()
    for j in range(1, N):
        os += (1.0, "Sz", j, "Sz", j + 1)
        os += (0.5, ("S+", j), ("S-", j + 1))
        os += (0.5, "S-", j, "S+", j + 1)
    for j in range(1, N + 1):
        os += (float(h[j - 1]), "Sz", j)
    H = T.CouplingModel(os, N)
    g, sn = T.default_graph_sitenodes(N)
    tree = TT._Tree(N, g, sn, T.find_eccentric_central_node(g, set(sn.values())))
    chan, W, site = TT._tree_operator(tree, H)
    otree = TO.Tree.build(N)
    ochan, oW, osite = TO.build_tree_operator(otree, TO.heisenberg_terms(N, 1.0, h))
    assert chan == ochan
    assert W.keys() == oW.keys()
    for n in W:
        assert [w[:3] for w in W[n]] == [w[:3] for w in oW[n]]
        assert np.allclose([w[3] for w in W[n]], [w[3] for w in oW[n]], rtol=0, atol=1e-15)
    for i in site:  # device layout [c][bra][ket] vs the oracle's [c][ket][bra]
        assert np.allclose(site[i], np.transpose(osite[i], (0, 2, 1)), rtol=0, atol=1e-15)


def test_opstrings_and_params_validation():
    os = T.OpStrings()
    with pytest.raises(ValueError, match="pairs"):
        os += (1.0, "Sz", 1, "Sz")
    os += (1.0, "Sz", 1, "Sz", 2, "Sz", 3)
    g, sn = T.default_graph_sitenodes(5)
    tree = TT._Tree(5, g, sn, T.find_eccentric_central_node(g, set(sn.values())))
    with pytest.raises(NotImplementedError):  # 3-body strings: MPS path only
        TT._tree_operator(tree, T.CouplingModel(os, 5))
    with pytest.raises(ValueError, match="Size mismatch"):
        T.make_optimize_params(maxdim=[16, 32], nsweeps=[1])
    p = T.make_optimize_params(maxdim=[128], nsweeps=[5], cutoff=1e-14, noise=1e-2, noisedecay=5,
                               disable_noise_after=2)
    assert p.noise == [1e-2] and p.noisedecay == [5.0] and p.disable_noise_after == [2]


def test_optimize_instances_host_paths():
    """The instance-parallel runner: an empty batch returns nothing, and an
    instance that fails its host-side validation (here a schedule whose maxdim
    and nsweeps lengths differ, make_optimize_params' size check) raises the
    same error in the caller, both in-process and from a spawned worker, after
    the pool has shut down."""
    from tennetlib_amd.batch import optimize_instances
    assert optimize_instances([], dict(maxdim=[16], nsweeps=[1])) == []
    bad = dict(maxdim=[16, 32], nsweeps=[1])
    spec = [(8, np.zeros(8), 0)]
    with pytest.raises(ValueError, match="Size mismatch"):
        optimize_instances(spec, bad, workers=1)
    with pytest.raises(ValueError, match="Size mismatch"):
        optimize_instances(spec * 2, bad, workers=2)


def test_map_instances_order_and_worker_death():
    """map_instances returns results in input order whatever order the workers
    finish in (submissions sorted by size), and a worker process that dies
    surfaces as BrokenProcessPool in the caller instead of a hang."""
    import operator
    import os
    from concurrent.futures.process import BrokenProcessPool
    from tennetlib_amd.batch import map_instances
    args = [(i, 3) for i in range(7)]
    want = [i * 3 for i in range(7)]
    assert map_instances(operator.mul, args, workers=1) == want
    assert map_instances(operator.mul, args, workers=3, size=lambda a: a[0]) == want
    with pytest.raises(BrokenProcessPool):
        map_instances(os._exit, [(3,), (3,)], workers=2)
