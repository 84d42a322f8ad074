"""GPU parity of the TTN path (tn_ttn C ABI) against the TTN oracle.

Tolerances: node products and environments 1e-12 relative (FP64, summation
order only); isometry moves keep the dense state to 1e-12; sweep energies of
the noiseless optimisation within 1e-10 of the oracle's (the QR gauges of the
two paths differ, the projected Krylov problem does not); ground energies
within 1e-9 of exact diagonalisation.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import tn_oracle_ttn as TO
from tests.test_ttn_oracle import E0_N12, dense_h, ttn_to_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import tennetlib_amd as T
    T.default_context()
    return T


def _model(T, N, h=None):
    os = T.OpStrings()
    for j in range(1, N):
        os += (1.0, "Sz", j, "Sz", j + 1)
        os += (0.5, "S+", j, "S-", j + 1)
        os += (0.5, "S-", j, "S+", j + 1)
    if h is not None:
        for j in range(1, N + 1):
            os += (float(h[j - 1]), "Sz", j)
    return T.CouplingModel(os, N), TO.heisenberg_terms(N, 1.0, h)


def _gpu_ttn(T, opsi: TO.TTNState):
    from tennetlib_amd import ttn as TT
    t = opsi.tree
    tree = TT._Tree(t.N, t.edges, t.sitenodes, t.root)
    return T.TTN(tree, {k: v.copy() for k, v in opsi.T.items()}, opsi.center)


def _as_oracle(psi, otree):
    return TO.TTNState(otree, {k: v.copy() for k, v in psi.tensors.items()}, psi.orthocenter)


@pytest.mark.parametrize("N,chi", [(12, 8), (23, 16), (32, 40)])
def test_ttn_product_matches_oracle_at_every_node(T, N, chi):
    rng = np.random.default_rng(N)
    h = rng.uniform(-2, 2, N)
    H, OH = _model(T, N, h)
    otree = TO.Tree.build(N)
    opsi = TO.random_ttn(otree, chi, seed=N)
    oenv = TO.StateEnvsTTN(opsi, OH)
    for node in TO.default_sweeppath(N, otree.edges):
        oenv.position(node)
        print(f"ttn product N={N} chi={chi} node={node}", flush=True)
        env = T.StateEnvsTTN(_gpu_ttn(T, oenv.psi), H)
        v = rng.standard_normal(oenv.psi.T[node].shape)
        want = oenv.env.product(node, v)
        got = T.product(env, v)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), node
        env.close()


def test_ttn_isometry_moves_keep_the_state(T):
    N = 12
    otree = TO.Tree.build(N)
    opsi = TO.random_ttn(otree, 8, seed=3)
    ref = ttn_to_dense(opsi)
    H, _ = _model(T, N)
    env = T.StateEnvsTTN(_gpu_ttn(T, opsi), H)
    for node in TO.default_sweeppath(N, otree.edges):
        print(f"ttn move to {node}", flush=True)
        T.position_(env, node, cutoff=-1.0)
        psi = _as_oracle(env.psi, otree)
        assert np.abs(ttn_to_dense(psi) - ref).max() < 1e-12
        # every other node is an isometry toward the center
        for n in otree.edges:
            if n == node:
                continue
            s = otree.slot_of(n, TO.shortest_path(otree.edges, n, node)[1])
            A = np.moveaxis(psi.T[n], s, -1)
            M = A.reshape(-1, A.shape[-1])
            assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < 1e-12, (node, n)
    env.close()


def test_ttn_truncating_move_matches_svd(T):
    """A move with cutoff > 0 / maxdim below the bond (ttn.jl:292-302): the
    kept weights and the truncated state equal the oracle's numpy SVD move."""
    N = 16
    otree = TO.Tree.build(N)
    opsi = TO.random_ttn(otree, 32, seed=6)
    a = otree.root
    b = sorted(otree.edges[a])[-1]
    want = opsi.copy()
    err_ref = TO.move_isometry(want, a, b, maxdim=10, cutoff=1e-14)
    H, _ = _model(T, N)
    env = T.StateEnvsTTN(_gpu_ttn(T, opsi), H)
    from tennetlib_amd import ttn as TT
    err = TT._move(env.ctx, env.h, env.tree, a, b, maxdim=10, cutoff=1e-14)
    env.orthocenter = b
    got = _as_oracle(env.psi, otree)
    assert abs(err - err_ref) < 1e-14
    assert got.T[b].shape == want.T[b].shape
    assert np.abs(ttn_to_dense(got) - ttn_to_dense(want)).max() < 1e-12
    env.close()


def test_ttn_optimize_matches_oracle(T):
    N = 16
    h = np.random.default_rng(7).uniform(-2, 2, N)
    H, OH = _model(T, N, h)
    otree = TO.Tree.build(N)
    opsi = TO.random_ttn(otree, 16, seed=2)
    oenv = TO.StateEnvsTTN(opsi, OH)
    path = TO.default_sweeppath(N, otree.edges)
    sw_ref = TO.optimize(oenv, TO.OptimizeParamsTTN.make([16], [3]), path)
    env = T.StateEnvsTTN(_gpu_ttn(T, opsi), H)
    sw = T.optimize_(env, T.make_optimize_params(maxdim=[16], nsweeps=[3]), T.default_sweeppath(env.psi),
                     outputlevel=0)
    print("ttn optimize", sw.energy, sw_ref.energy, flush=True)
    assert sw.maxchi == sw_ref.maxchi
    assert np.allclose(sw.energy, sw_ref.energy, rtol=0, atol=1e-10)
    env.close()


def test_ttn_reference_settings_reach_ed(T):
    """test/test_TTN.jl:27-38 settings (default_randomTTN(sites, 12), maxdim
    128, 5 sweeps, cutoff 1e-14, noise 1e-2, noisedecay 5,
    disable_noise_after 2) on the N = 12 Heisenberg chain."""
    N = 12
    H, _ = _model(T, N)
    psi0 = T.default_randomTTN(N, 12, seed=1)
    p = T.make_optimize_params(maxdim=[128], nsweeps=[5], cutoff=1e-14, noise=1e-2, noisedecay=5,
                               disable_noise_after=2)
    en, psi = T.optimize(psi0, H, p, T.default_sweeppath(psi0), outputlevel=0)
    print("ttn reference settings", en, flush=True)
    assert abs(en - E0_N12) < 1e-9
    assert psi.maxlinkdim() > 12


def test_ttn_random_field_instance_reaches_ed(T):
    N = 16
    h = np.random.default_rng(11).uniform(-2, 2, N)
    H, OH = _model(T, N, h)
    e0 = spla.eigsh(dense_h(OH), k=1, which="SA", tol=1e-12)[0][0]
    psi0 = T.default_randomTTN(N, 64, seed=2)
    en, psi = T.optimize(psi0, H, T.make_optimize_params(maxdim=[64], nsweeps=[4]), T.default_sweeppath(psi0),
                         outputlevel=0)
    print("ttn random field", en, e0, flush=True)
    assert abs(en - e0) < 1e-9


def test_ttn_concurrent_contexts_match_sequential(T):
    """Regression: two host threads optimising different instances on their
    own contexts (the library serialises its entry points) reproduce the
    sequential results bit for bit."""
    import threading
    specs = [(16, 3), (20, 4)]
    params = T.make_optimize_params(maxdim=[64], nsweeps=[3], cutoff=1e-14, noise=1e-2, noisedecay=5,
                                    disable_noise_after=2)

    def run(N, seed, ctx, out):
        h = np.random.default_rng(seed).uniform(-2, 2, N)
        H, _ = _model(T, N, h)
        psi0 = T.default_randomTTN(N, 12, seed=seed, ctx=ctx)
        env = T.StateEnvsTTN(psi0, H, ctx=ctx)
        sw = T.optimize_(env, params, T.default_sweeppath(psi0), outputlevel=0, seed=seed)
        env.close()
        out[(N, seed)] = (list(sw.energy), list(sw.maxchi))

    seq, par = {}, {}
    ctx0 = T.Context(0)
    for N, seed in specs:
        run(N, seed, ctx0, seq)
    ctxs = [T.Context(0) for _ in specs]
    ths = [threading.Thread(target=run, args=(N, seed, c, par)) for (N, seed), c in zip(specs, ctxs)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    print("ttn concurrent", seq, par, flush=True)
    assert par == seq


def test_ttn_errors(T):
    N = 8
    H, _ = _model(T, N)
    psi0 = T.default_randomTTN(N, 4)
    env = T.StateEnvsTTN(psi0, H)
    with pytest.raises(ValueError, match="index mismatch"):
        T.product(env, np.zeros(3))
    with pytest.raises(ValueError, match="ill-defined"):
        T.position_(env, (9, 9))
    with pytest.raises(ValueError, match="expand_numiter"):
        T.fullsweep_(env, T.default_sweeppath(psi0), T.eig_solver, T.SweepDataTTN(), noise=1e-2, expand_numiter=1)
    env.close()


def test_ttn_projector_product_matches_oracle(T):
    """Node products of StateEnvsTTN(psi, H, Ms; weight) (LinkProjTTN) at every
    node, two projector TTNs of other bond dimensions."""
    N, chi = 12, 8
    rng = np.random.default_rng(41)
    H, OH = _model(T, N, rng.uniform(-1, 1, N))
    otree = TO.Tree.build(N)
    oMs = [TO.random_ttn(otree, chi + 2 + k, seed=50 + k) for k in range(2)]
    opsi = TO.random_ttn(otree, chi, seed=42)
    oenv = TO.StateEnvsTTN(opsi, OH, oMs, weight=2.5)
    for node in TO.default_sweeppath(N, otree.edges):
        oenv.position(node)
        print(f"ttn projector product node={node}", flush=True)
        env = T.StateEnvsTTN(_gpu_ttn(T, oenv.psi), H, [_gpu_ttn(T, M) for M in oMs], weight=2.5)
        v = rng.standard_normal(oenv.psi.T[node].shape)
        want = oenv.env.product(node, v)
        got = T.product(env, v)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), node
        env.close()


def test_ttn_projector_optimize_matches_oracle(T):
    N = 12
    H, OH = _model(T, N)
    otree = TO.Tree.build(N)
    path = TO.default_sweeppath(N, otree.edges)
    ogs = TO.StateEnvsTTN(TO.random_ttn(otree, 16, seed=3), OH)
    TO.optimize(ogs, TO.OptimizeParamsTTN.make([16], [3]), path)
    opsi = TO.random_ttn(otree, 16, seed=4)
    oenv = TO.StateEnvsTTN(opsi, OH, [ogs.psi], weight=10.0)
    sw_ref = TO.optimize(oenv, TO.OptimizeParamsTTN.make([16], [3]), path)
    env = T.StateEnvsTTN(_gpu_ttn(T, opsi), H, [_gpu_ttn(T, ogs.psi)], weight=10.0)
    sw = T.optimize_(env, T.make_optimize_params(maxdim=[16], nsweeps=[3]), T.default_sweeppath(env.psi),
                     outputlevel=0)
    print("ttn projector optimize", sw.energy, sw_ref.energy, flush=True)
    assert sw.maxchi == sw_ref.maxchi
    assert np.allclose(sw.energy, sw_ref.energy, rtol=0, atol=1e-10)
    env.close()


def test_ttn_excited_state_reference_settings(T):
    """test_TTN.jl:40-64 (dense): ground state from the reference settings,
    then optimize(psi0, H, [psi_gr], params, sweeppath; weight = 10) with
    maxdim 128, 10 sweeps, noise 1e-2 (decay 5, off after 6): the lowest
    triplet E1 = -4.86114793703639 (ED), orthogonal to the ground state."""
    N = 12
    H, _ = _model(T, N)
    psi0 = T.default_randomTTN(N, 12, seed=1)
    p = T.make_optimize_params(maxdim=[128], nsweeps=[5], cutoff=1e-14, noise=1e-2, noisedecay=5,
                               disable_noise_after=2)
    e0, gs = T.optimize(psi0, H, p, T.default_sweeppath(psi0), outputlevel=0)
    p1 = T.make_optimize_params(maxdim=[128], nsweeps=[10], cutoff=1e-14, noise=1e-2, noisedecay=5,
                                disable_noise_after=6)
    e1, ex = T.optimize(psi0, H, [gs], p1, T.default_sweeppath(psi0), weight=10.0, outputlevel=0)
    print("ttn excited", e0, e1, flush=True)
    assert abs(e0 - E0_N12) < 1e-9
    assert abs(e1 - (-4.86114793703639)) < 1e-8
    otree = TO.Tree.build(N)
    ov = ttn_to_dense(_as_oracle(gs, otree)) @ ttn_to_dense(_as_oracle(ex, otree))
    assert abs(ov) < 1e-5
    with pytest.raises(ValueError, match="should be > 0.0"):
        T.StateEnvsTTN(psi0, H, [gs], weight=0.0)


def test_ttn_instances_in_worker_processes_match_sequential(T):
    """Instance-parallel runs (optimize_instances: spawned worker processes
    sharing the GPU, each with its own library context) against the same
    instances run one after another in this process, for the noiseless and
    the noisy schedule (noise 1e-2, decay 5, off after 2; cutoff 1e-14; 5
    sweeps), maxdim 64.  Noiseless (4 sweeps at the initial bond dimension
    12): the same energies to 1e-12 per site and the same bond dimensions.
    Noisy: energies to 1e-9 per site (the cutoff may split a rounding-level
    spectrum differently, so bond dimensions are not compared); the noisy
    N = 12 instance meets exact diagonalisation to 1e-9."""
    from tennetlib_amd.batch import ground_state, optimize_instances
    rng = np.random.default_rng(11)
    specs = [(N, rng.uniform(-2, 2, N), s) for s, N in enumerate([12, 16, 20, 24])]
    quiet = dict(maxdim=[64], nsweeps=[4], cutoff=0.0)
    noisy = dict(maxdim=[64], nsweeps=[5], cutoff=1e-14, noise=1e-2, noisedecay=5, disable_noise_after=2)
    runs = {}
    for name, pkw, tol in (("noiseless", quiet, 1e-12), ("noisy", noisy, 1e-9)):
        print(f"ttn instances {name}: sequential", flush=True)
        seq = [ground_state(N, h, s, pkw) for N, h, s in specs]
        print(f"ttn instances {name}: 2 worker processes", flush=True)
        par = optimize_instances(specs, pkw, workers=2)
        dE = [abs(a[0][-1] - b[0][-1]) / N for a, b, (N, _, _) in zip(seq, par, specs)]
        exact = all(a == b for a, b in zip(seq, par))
        print(f"ttn instances {name}: max |dE/N| {max(dE):.3e}, identical {exact}", flush=True)
        assert max(dE) <= tol
        if name == "noiseless":
            assert [a[1] for a in seq] == [b[1] for b in par]
        runs[name] = par
    # the noisy schedule grows the bonds from 12; at N = 12 bond dimension 64
    # holds the ground state exactly: exact diagonalisation
    N0, h0, _ = specs[0]
    e0 = spla.eigsh(dense_h(TO.heisenberg_terms(N0, 1.0, h0)), k=1, which="SA", tol=1e-12)[0][0]
    assert abs(runs["noisy"][0][0][-1] - e0) <= 1e-9
