"""GPU parity of StateEnvs over a CouplingModel (ProjCouplingModel) and over
a list of MPOs (ProjMPOSum2) against the oracle's per-id restatement.

Tolerances: products 1e-12 relative (FP64, summation order only); noisy
bond replacement: kept spectrum 1e-10 relative, truncation error 1e-12, the
product of the two new sites 1e-10 relative; sweep energies as in the MPO
noisy-DMRG parity test: the noisy sweeps' two-iteration Lanczos Ritz values
depend on how the cutoff splits near-degenerate density-matrix multiplets.
With every density matrix perturbed by 2e-16 * max|rho| before its eigh, the
oracle's first-sweep energy moves by up to 6e-12 for the CouplingModel chain
and by up to 1.9e-9 for the MPO sum (later sweeps < 2e-14): 1e-9 on both
noisy sweeps of the former, 1e-8 on the first of the latter, 1e-11 after
them; ground energies 1e-9 of exact diagonalisation.
"""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import tn_oracle as O
from oracle import tn_oracle_cm as CM
from tests.test_oracle import _mixed_canonical
from tests.test_oracle_cm import kron_term, mixed_model_terms

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


def host_model(T, terms, N):
    os = T.OpStrings()
    for c, ops in terms:
        os += (c, *[(o, i) for o, i in ops])
    return T.CouplingModel(os, N)


def heis_terms(N, h):
    t = []
    for j in range(1, N):
        t += [(1.0, [("Sz", j), ("Sz", j + 1)]), (0.5, [("S+", j), ("S-", j + 1)]),
              (0.5, [("S-", j), ("S+", j + 1)])]
    return t + [(float(h[j - 1]), [("Sz", j)]) for j in range(1, N + 1)]


@pytest.mark.parametrize("cplx", [False, True])
def test_coupling_model_product_matches_oracle(T, cplx):
    """Per-id product of ProjCouplingModel (1-, 2- and 3-site strings, a
    same-site product, fields) on a mixed-canonical MPS at every position
    checked, one and two sites."""
    N = 8
    terms = mixed_model_terms(N, 5)
    H = host_model(T, terms, N)
    ref_model = CM.CMModel.from_terms(terms, N)
    A0 = rand_mps(N, [2, 4, 8, 16, 8, 4, 2], seed=31, cplx=cplx)
    rng = np.random.default_rng(32)
    for nsite, positions in ((2, [1, 4, 7]), (1, [1, 5, 8])):
        for p in positions:
            print(f"cm product cplx={cplx} nsite={nsite} pos={p}", flush=True)
            A = _mixed_canonical(A0, p, nsite)
            env = T.StateEnvs(A, H)
            T.set_nsite_(env, nsite)
            T.position_(env, p)
            ref = CM.ProjCM(ref_model)
            ref.nsite = nsite
            ref.position(A, p)
            dims, _ = env.phi_dims()
            shape = dims if nsite == 2 else (dims[0], dims[1], dims[3])
            v = rng.standard_normal(shape)
            if cplx:
                v = v + 1j * rng.standard_normal(shape)
            want = ref.product(v)
            got = T.product(env, v).reshape(want.shape)
            assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (nsite, p)
            env.close()


def test_mpo_sum_product_matches_oracle(T):
    N = 7
    W1 = O.xxz_mpo(N, 1.0, 0.6)
    W2 = O.xxz_mpo(N, 0.0, 0.0, np.linspace(-0.8, 0.9, N))
    A = rand_mps(N, [2, 4, 8, 8, 4, 2], seed=41)
    env = T.StateEnvs(A, [W1, W2])
    ref = CM.ProjSum([W1, W2])
    rng = np.random.default_rng(42)
    for p in (1, 3, 6, 2):
        print(f"mpo sum product pos={p}", flush=True)
        T.position_(env, p)
        ref.position(A, p)
        dims, _ = env.phi_dims()
        v = rng.standard_normal(dims)
        want = ref.product(v)
        got = T.product(env, v).reshape(want.shape)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), p
    env.close()


def _noisy_replacebond(T, env, b, phi, ortho, noise, maxdim):
    ctx = env.ctx
    dphi = T.DeviceArray.from_numpy(ctx, phi)
    terr, nk = C.c_double(), C.c_long()
    eigs = (C.c_double * 64)()
    ctx.check(ctx.lib.tn_replacebond(env.h, b, dphi.ptr, 1 if ortho == "left" else 0, maxdim, 1, 1e-12, noise, 1,
                                     C.byref(terr), eigs, 64, C.byref(nk)))
    return np.array(eigs[:nk.value]), terr.value


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ortho", ["left", "right"])
@pytest.mark.parametrize("kind", ["couplingmodel", "mposum"])
def test_noisy_replacebond_matches_oracle(T, kind, ortho, cplx):
    """replacebond! with eigen_perturbation = noise * noiseterm(PH, phi, ortho)
    for ProjCouplingModel (sum over the ids acting on the kept side) and
    ProjMPOSum2 (sum of the ProjMPO noise terms)."""
    N, noise, maxdim = 8, 1e-2, 6
    h = np.linspace(-0.6, 0.7, N)
    A0 = rand_mps(N, [2, 4, 8, 12, 8, 4, 2], seed=51, cplx=cplx)
    for b in (1, 4, 7):
        print(f"noisy replacebond kind={kind} ortho={ortho} cplx={cplx} b={b}", flush=True)
        A = _mixed_canonical(A0, b, 2)
        if kind == "couplingmodel":
            terms = mixed_model_terms(N, 6)
            env = T.StateEnvs(A, host_model(T, terms, N))
            ref = CM.ProjCM(CM.CMModel.from_terms(terms, N))
        else:
            mpos = [O.xxz_mpo(N, 1.0, 0.6), O.xxz_mpo(N, 0.0, 0.0, h)]
            env = T.StateEnvs(A, mpos)
            ref = CM.ProjSum(mpos, dtype=np.complex128 if cplx else np.float64)
        T.position_(env, b)
        ref.position(A, b)
        phi = np.tensordot(A[b - 1], A[b], axes=([2], [0]))
        phi = phi / np.linalg.norm(phi)
        drho = noise * O.noiseterm(ref, phi, ortho)
        Aref = [None, None]
        sp = O.replacebond_perturbed(Aref, 0, phi, maxdim, 1, 1e-12, ortho, True, drho)
        eigs, terr = _noisy_replacebond(T, env, b, phi, ortho, noise, maxdim)
        assert len(eigs) == len(sp.eigs)
        assert np.allclose(eigs, sp.eigs, rtol=1e-10, atol=1e-14)
        assert abs(terr - sp.truncerr) < 1e-12
        got = env.psi
        prod = np.tensordot(got[b - 1], got[b], axes=([2], [0]))
        want = np.tensordot(Aref[0], Aref[1], axes=([2], [0]))
        assert np.abs(prod - want).max() < 1e-10 * np.abs(want).max()
        env.close()


def test_coupling_model_noisy_dmrg_matches_oracle(T):
    """The reference's noisy schedule (noise 1e-3, noisedecay 2,
    disable_noise_after 2) on a random-field Heisenberg chain given as
    OpStrings: sweep energies follow the oracle's ProjCouplingModel DMRG and
    reach exact diagonalisation.  Fields within +-0.3 keep the ground state
    in the Neel start's Sz = 0 sector: with stronger fields the sweeps must
    change sector, and without quantum numbers the wave function reaches the
    new sector's basis states (which the noise term adds) only through
    rounding, so the sweep at which it happens follows rounding."""
    N = 12
    h = np.random.default_rng(61).uniform(-0.3, 0.3, N)
    terms = heis_terms(N, h)
    kw = dict(cutoff=1e-12, noise=1e-3, noisedecay=2.0, disable_noise_after=2)
    ref = O.StateEnvs(O.neel_mps(N), CM.CMModel.from_terms(terms, N))
    sw_ref = O.dmrg(ref, O.DMRGParams.make([64], [6], **kw))
    env = T.StateEnvs(O.neel_mps(N), host_model(T, terms, N))
    sw = T.dmrg_(env, T.make_dmrg_params(maxdim=[64], nsweeps=[6], **kw), 2, outputlevel=0)
    print("cm noisy dmrg", sw.energy, sw_ref.energy, flush=True)
    assert np.allclose(sw.energy[:2], sw_ref.energy[:2], rtol=0, atol=1e-9)
    assert np.allclose(sw.energy[2:], sw_ref.energy[2:], rtol=0, atol=1e-11)
    assert sw.maxchi == sw_ref.maxchi
    Hd = sum(kron_term(N, c, ops) for c, ops in terms)
    e0 = spla.eigsh(Hd, k=1, which="SA", tol=1e-12)[0][0]
    assert abs(sw.energy[-1] - e0) < 1e-9


def test_mpo_sum_noisy_dmrg_matches_oracle(T):
    """As above for StateEnvs(psi, [H_exchange, H_field]) (ProjMPOSum2)."""
    N = 12
    h = np.random.default_rng(62).uniform(-0.3, 0.3, N)
    mpos = [O.heisenberg_mpo(N), O.xxz_mpo(N, 0.0, 0.0, h)]
    kw = dict(cutoff=1e-12, noise=1e-3, noisedecay=2.0, disable_noise_after=2)
    ref = O.StateEnvs(O.neel_mps(N), mpos)
    sw_ref = O.dmrg(ref, O.DMRGParams.make([64], [6], **kw))
    env = T.StateEnvs(O.neel_mps(N), mpos)
    sw = T.dmrg_(env, T.make_dmrg_params(maxdim=[64], nsweeps=[6], **kw), 2, outputlevel=0)
    print("mpo sum noisy dmrg", sw.energy, sw_ref.energy, flush=True)
    assert abs(sw.energy[0] - sw_ref.energy[0]) < 1e-8
    assert abs(sw.energy[1] - sw_ref.energy[1]) < 1e-9
    assert np.allclose(sw.energy[2:], sw_ref.energy[2:], rtol=0, atol=1e-11)
    e0 = spla.eigsh(O.mpo_to_dense(O.xxz_mpo(N, 1.0, 1.0, h)), k=1, which="SA", tol=1e-12)[0][0]
    assert abs(sw.energy[-1] - e0) < 1e-9


def test_coupling_model_excited_state(T):
    """StateEnvs(psi, H::CouplingModel, Ms; weight) (ProjCouplingModel_MPS):
    the N = 12 Heisenberg chain's first excited level by exact
    diagonalisation, E1 = -4.86114793703639."""
    N = 12
    H = host_model(T, heis_terms(N, np.zeros(N)), N)
    e0, gs = T.dmrg(O.neel_mps(N), H, T.make_dmrg_params(maxdim=[64], nsweeps=[4], cutoff=1e-14), 2,
                    outputlevel=0)
    e1, ex = T.dmrg(O.random_product_mps(N, 2, seed=3), H, [gs], T.make_dmrg_params(maxdim=[64], nsweeps=[6],
                                                                                    cutoff=1e-14), 2,
                    weight=10.0, outputlevel=0)
    print("cm excited", e0, e1, flush=True)
    assert abs(e0 - (-5.14209063284054)) < 1e-9
    assert abs(e1 - (-4.86114793703639)) < 1e-8
    with pytest.raises(ValueError, match="should be > 0.0"):
        T.StateEnvs(O.neel_mps(N), H, [gs], weight=0.0)
