"""GPU parity of one-site updates, one-site / dynamic sweeps and the global
subspace expansion (tn_update_one_site, tn_env_krylov_extend, nsite = 0)
against the oracle (oracle/tn_oracle_dyn.py).

Tolerances: zero-site products 1e-12 relative (FP64); one-site DMRG sweep
energies 1e-10 of the oracle's (Ritz values of the same projected problem);
noisy one-site sweeps 1e-8 on the first sweep and 1e-9 after it (the
noise-perturbed factorisation keeps near-degenerate density-matrix multiplets
whose split follows rounding: the oracle itself moves the first sweep's
energy by up to 4.3e-9 when every density matrix is perturbed by
2e-16 * max|rho| before its eigh, later sweeps by < 1e-14);
time evolution: overlap with exact evolution 1e-8, energy conservation 1e-9;
subspace expansion: bond dimensions equal to the oracle's, the right-basis
projectors within 1e-8 and the state kept to 2e-9 max|v0|. Perturbing every
matrix the oracle diagonalises by a random Hermitian 2e-16 * max|.| moves its own
reconstruction by up to 4.7e-10 max|v0| on the real case (five perturbations;
complex: 5e-12), so the state is fixed only to that level by the algorithm: the
basis added from eigenvectors of (1 - B^H B) G (1 - B^H B) near the 1e-7 cutoff
carries rounding components along B of order eps ||G|| / lambda.
"""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import tn_oracle as O
from oracle import tn_oracle_cm as CM
from oracle import tn_oracle_dyn as D
from tests.test_oracle import _right_iso, _sz_expect, xxz_dense
from tests.test_oracle_dyn import _full_random_mps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import tennetlib_amd as T
    T.default_context()
    return T


def rand_mps(N, chis, seed, cplx=False):
    rng = np.random.default_rng(seed)
    dims = [1] + list(chis) + [1]
    A = []
    for k in range(N):
        a = rng.standard_normal((dims[k], 2, dims[k + 1]))
        if cplx:
            a = a + 1j * rng.standard_normal(a.shape)
        A.append(a / np.linalg.norm(a))
    return A


@pytest.mark.parametrize("cplx", [False, True])
def test_zero_site_product_matches_oracle(T, cplx):
    N = 7
    Ws = O.xxz_mpo(N, 1.0, 0.7, np.linspace(-0.5, 0.6, N))
    A = rand_mps(N, [2, 4, 9, 9, 4, 2], seed=171, cplx=cplx)
    env = T.StateEnvs(A, Ws)
    ref = O.Envs(Ws, dtype=np.complex128 if cplx else np.float64)
    T.set_nsite_(env, 0)
    ref.nsite = 0
    rng = np.random.default_rng(172)
    for p in (2, 4, 7, 3):
        print(f"zero-site product cplx={cplx} pos={p}", flush=True)
        T.position_(env, p)
        ref.position(A, p)
        shape = (A[p - 2].shape[2], A[p - 1].shape[0])
        v = rng.standard_normal(shape)
        if cplx:
            v = v + 1j * rng.standard_normal(shape)
        want = ref.product(v)
        got = T.product(env, v).reshape(want.shape)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), p
    env.close()


def test_one_site_dmrg_matches_oracle(T):
    N = 10
    h = np.random.default_rng(173).uniform(-0.4, 0.4, N)
    Ws = O.xxz_mpo(N, 1.0, 1.0, h)
    A0 = _full_random_mps(N, 174)
    ref = O.StateEnvs(A0, Ws)
    sw_ref = O.SweepData()
    for _ in range(3):
        D.fullsweep(ref, sw_ref, 1, maxdim=64, cutoff=1e-14)
    env = T.StateEnvs(A0, Ws)
    sw = T.SweepData()
    for _ in range(3):
        T.fullsweep_(env, T.eig_solver, 1, sw, maxdim=64, cutoff=1e-14, outputlevel=0)
    print("dmrg1", sw.energy, sw_ref.energy, flush=True)
    assert np.allclose(sw.energy, sw_ref.energy, rtol=0, atol=1e-10)
    assert sw.maxchi == sw_ref.maxchi
    e0 = spla.eigsh(O.mpo_to_dense(Ws), k=1, which="SA")[0][0]
    assert abs(sw.energy[-1] - e0) < 1e-9
    # dmrg1 entry point
    e1, _ = T.dmrg1(A0, Ws, T.make_dmrg_params(maxdim=[64], nsweeps=[3], cutoff=1e-14), outputlevel=0)
    assert abs(e1 - sw.energy[-1]) < 1e-10


def test_one_site_noisy_dmrg_matches_oracle(T):
    """One-site DMRG from Neel grows its bonds only through the noise term
    (the two-site density-matrix factorisation of update_site.jl:139-159)."""
    N = 10
    Ws = O.heisenberg_mpo(N)
    ref = O.StateEnvs(O.neel_mps(N), Ws)
    sw_ref = O.SweepData()
    env = T.StateEnvs(O.neel_mps(N), Ws)
    sw = T.SweepData()
    noise = 1e-3
    for k in range(7):
        nz = noise if k < 4 else 0.0
        D.fullsweep(ref, sw_ref, 1, maxdim=64, cutoff=1e-14, noise=nz)
        T.fullsweep_(env, T.eig_solver, 1, sw, maxdim=64, cutoff=1e-14, noise=nz, outputlevel=0)
        noise /= 2
    print("dmrg1 noise", sw.energy, sw_ref.energy, flush=True)
    assert abs(sw.energy[0] - sw_ref.energy[0]) < 1e-8
    assert np.allclose(sw.energy[1:], sw_ref.energy[1:], rtol=0, atol=1e-9)
    assert sw.maxchi[-1] > 8
    e0 = spla.eigsh(O.mpo_to_dense(Ws), k=1, which="SA")[0][0]
    assert abs(sw.energy[-1] - e0) < 1e-9


def test_tdvp1_full_bond_dimension_matches_free_fermions(T):
    N, dt, steps = 8, 0.05, 10
    Ws = O.xxz_mpo(N, 1.0, 0.0)
    chis = [1] + [min(2 ** k, 2 ** (N - k)) for k in range(1, N)] + [1]
    A = []
    for k in range(N):
        t = np.zeros((chis[k], 2, chis[k + 1]), dtype=complex)
        t[0, k % 2, 0] = 1.0
        A.append(t)
    eng = T.TDVPEngine(A, Ws)
    for _ in range(steps):
        T.tdvpsweep_(eng, -1j * dt, 1, maxdim=64, cutoff=-1.0, outputlevel=0)
    hsp = 0.5 * (np.eye(N, k=1) + np.eye(N, k=-1))
    Ut = sla.expm(-1j * hsp * dt * steps)
    occ = [i for i in range(N) if i % 2 == 0]
    n = (np.abs(Ut[:, occ]) ** 2).sum(1)
    got = _sz_expect(eng.sysenv.psi)
    print("tdvp1 sz", np.abs(got - (n - 0.5)).max(), flush=True)
    assert np.abs(got - (n - 0.5)).max() < 1e-8
    assert abs(eng.swdata.energy[-1]) < 1e-8


def _right_projectors(A):
    N = len(A)
    out = []
    for j in range(2, N + 1):
        Rb = _right_iso(A, j - 1)
        out.append(Rb.T @ Rb.conj())
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_krylov_extend_matches_oracle(T, cplx):
    N = 8
    Ws = O.heisenberg_mpo(N)
    A0 = D.orthogonalize(rand_mps(N, [2, 3, 3, 3, 3, 3, 2], seed=175, cplx=cplx), 1)
    A0[0] = A0[0] / np.linalg.norm(A0[0])
    ref = O.StateEnvs(A0, Ws)
    D.krylov_extend(ref)
    env = T.StateEnvs(A0, Ws)
    T.orthogonalize_(env, 1)
    T.krylov_extend_(env, outputlevel=0)
    got = env.psi
    print("gse dims", [a.shape[2] for a in got], [a.shape[2] for a in ref.psi], flush=True)
    assert [a.shape[2] for a in got] == [a.shape[2] for a in ref.psi]
    v0 = O.mps_to_dense(A0).reshape(-1)
    assert np.abs(O.mps_to_dense(got).reshape(-1) - v0).max() < 2e-9 * np.abs(v0).max()
    for Pg, Pr in zip(_right_projectors(got), _right_projectors(ref.psi)):
        assert np.abs(Pg - Pr).max() < 1e-8
    assert env.orthocenter == 1
    env.close()


@pytest.mark.parametrize("kind", ["mpo", "couplingmodel"])
def test_dynamic_tdvp_reference_settings(T, kind):
    """test_MPS_TDVP.jl:47-57: 10 tdvpsweep!(engine, -0.01im; nsite="dynamic",
    maxdim = 20, cutoff = 1e-12, extendat = 5) on the N = 12 Heisenberg
    chain from Neel, with an MPO (subspace expansion + one-site sweeps at
    sweeps 1, 5, 10) and with a CouplingModel (two-site sweeps there)."""
    N = 12
    terms = []
    for j in range(1, N):
        terms += [(1.0, [("Sz", j), ("Sz", j + 1)]), (0.5, [("S+", j), ("S-", j + 1)]),
                  (0.5, [("S-", j), ("S+", j + 1)])]
    if kind == "mpo":
        H, Href = O.heisenberg_mpo(N), O.heisenberg_mpo(N)
    else:
        os = T.OpStrings()
        for c, ops in terms:
            os += (c, *[(o, i) for o, i in ops])
        H, Href = T.CouplingModel(os, N), CM.CMModel.from_terms(terms, N)
    psi0 = [a.astype(complex) for a in O.neel_mps(N)]
    eng = T.TDVPEngine(psi0, H)
    ref = O.StateEnvs(psi0, Href)
    sw_ref = O.SweepData()
    for k in range(10):
        print(f"dynamic tdvp {kind} sweep {k + 1}", flush=True)
        T.tdvpsweep_(eng, -0.01j, "dynamic", maxdim=20, cutoff=1e-12, extendat=5, outputlevel=0)
        D.tdvpsweep(ref, sw_ref, -0.01j, nsite="dynamic", maxdim=20, cutoff=1e-12, extendat=5)
    sw = eng.swdata
    print("dynamic tdvp energies", sw.energy, sw_ref.energy, sw.maxchi, sw_ref.maxchi, flush=True)
    assert sw.maxchi == sw_ref.maxchi
    assert np.allclose(sw.energy, sw_ref.energy, rtol=0, atol=1e-9)
    assert np.allclose(sw.energy, -2.75, atol=1e-9)
    v0 = O.mps_to_dense(psi0).reshape(-1)
    vt = spla.expm_multiply(-1j * 0.1 * xxz_dense(N, sparse=True), v0)
    got = O.mps_to_dense(eng.sysenv.psi).reshape(-1)
    assert abs(abs(np.vdot(got, vt)) - 1.0) < 1e-8
    assert abs(eng.abstime - 0.1) < 1e-12


def test_zero_site_product_with_projectors_matches_oracle(T):
    """H + w sum_k |M_k><M_k| on the bond matrix (nsite = 0) of a complex
    MPS: the projectors' local factor is conj(L_M) conj(R_M)^T."""
    N, w = 6, 2.5
    Ws = O.xxz_mpo(N, 1.0, 0.7, np.linspace(-0.4, 0.6, N))
    A = rand_mps(N, [2, 4, 8, 4, 2], seed=181, cplx=True)
    Ms = [rand_mps(N, [2, 3, 5, 3, 2], seed=182 + k, cplx=True) for k in range(2)]
    env = T.StateEnvs(A, Ws, Ms, weight=w)
    ref = O.Envs(Ws, dtype=np.complex128, Ms=Ms, weight=w)
    T.set_nsite_(env, 0)
    ref.nsite = 0
    rng = np.random.default_rng(183)
    for p in (2, 4, 6, 3):
        print(f"zero-site projector product pos={p}", flush=True)
        T.position_(env, p)
        ref.position(A, p)
        shape = (A[p - 2].shape[2], A[p - 1].shape[0])
        v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        want = ref.product(v)
        got = T.product(env, v).reshape(want.shape)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), p
    env.close()
