"""Host lowering of CouplingModel and MPO sums (tennetlib_amd.couplingmodel)
against the oracle -- CPU only, no GPU calls.

The StateEnvs path runs a CouplingModel (or a list of MPOs) as one MPO whose
bond channels are the term ids; the noise factor of that MPO restricted to
the channels with an operator on the kept side must give the reference's
per-id noise term.  Checked here with the oracle's own MPO noise factor, at
every bond and on both sides, on a complex mixed-canonical MPS.
"""

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle import tn_oracle_cm as CM
from tests.test_oracle import _complex_mps, _mixed_canonical
from tests.test_oracle_cm import kron_term, mixed_model_terms
import tennetlib_amd as T
from tennetlib_amd import couplingmodel as HC
from tennetlib_amd import ttn as TT


def host_model(terms, N, **kw):
    os = T.OpStrings()
    for c, ops in terms:
        os += (c, *[(o, i) for o, i in ops])
    return T.CouplingModel(os, N, **kw)


def _left_grams(fs):
    """Per internal link, the Gram matrix of the partial string left of it
    (invariant under rotations of the link)."""
    out, P = [], np.ones((1, 1))  # P[(ops), link]
    for F in fs[:-1]:
        l, d1, d2, r = F.shape
        P = np.einsum("Xl,lstr->Xstr", P, F).reshape(-1, r)
        out.append(P @ P.T)
    return out


def test_ids_match_oracle():
    N = 7
    terms = mixed_model_terms(N, 3)
    host = host_model(terms, N).ids()
    ref = CM.CMModel.from_terms(terms, N).ids
    assert sorted(p for p, _ in host) == sorted(p for p, _ in ref)
    rd = dict(ref)
    for p, fs in host:
        assert len(fs) == len(rd[p])
        for Gh, Gr in zip(_left_grams(fs), _left_grams(rd[p])):
            assert np.abs(Gh - Gr).max() < 1e-14
    Hd = sum(kron_term(N, c, ops) for c, ops in terms)
    Ws, _, _ = HC.coupling_mpo(host, N, [2] * N)
    assert np.abs(O.mpo_to_dense(Ws) - Hd).max() < 1e-14


def _restricted_noise(Ws, sl, sr, A, p, phi, ortho):
    env = O.Envs(Ws, dtype=np.complex128)
    env.nsite = 2
    env.position(A, p)
    a, s, t, c = phi.shape
    X = O.noiseterm_factor(env, phi, ortho)
    if ortho == "left":
        X = X[:, sl[p] * t * c:]
        return X @ X.conj().T
    w = Ws[p].shape[0]
    X = X.reshape(a, s, w, t * c)[:, :, :w - sr[p]].reshape(-1, t * c)
    return X.T @ X.conj()


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_restricted_mpo_noise_equals_per_id_noise(p):
    N, d = 7, 2
    terms = mixed_model_terms(N, 4)
    Ws, sl, sr = HC.coupling_mpo(host_model(terms, N).ids(), N, [2] * N)
    ref = CM.CMModel.from_terms(terms, N)
    A = _mixed_canonical(_complex_mps(N, d, [2, 4, 6, 6, 4, 2], seed=110 + p), p, 2)
    env = CM.ProjCM(ref)
    env.position(A, p)
    rng = np.random.default_rng(p)
    shape = (A[p - 1].shape[0], d, d, A[p].shape[2])
    phi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for ortho in ("left", "right"):
        want = env.noiseterm(phi, ortho)
        got = _restricted_noise(Ws, sl, sr, A, p, phi, ortho)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), ortho


def test_mpo_sum_and_mpo_terms_lowering():
    N = 6
    h = np.linspace(-0.5, 0.7, N)
    W1, W2 = O.xxz_mpo(N, 1.0, 0.6), O.xxz_mpo(N, 0.0, 0.0, h)
    Ws, sl, sr = HC.coupling_mpo(HC.mpo_sum_ids([W1, W2]), N, [2] * N)
    assert not any(sl) and not any(sr)
    assert [W.shape[3] for W in Ws[:-1]] == [W1[k].shape[3] + W2[k].shape[3] for k in range(N - 1)]
    assert np.abs(O.mpo_to_dense(Ws) - O.mpo_to_dense(W1) - O.mpo_to_dense(W2)).max() < 1e-14
    # CouplingModel(os, mpo) and CouplingModel(mpo1, mpo2)
    os = T.OpStrings()
    os += (0.3, "Sx", 2, "Sx", 5)
    M = T.CouplingModel(os, W1)
    Ws, _, _ = HC.coupling_mpo(M.ids(), N, [2] * N)
    want = O.mpo_to_dense(W1) + kron_term(N, 0.3, [("Sx", 2), ("Sx", 5)])
    assert np.abs(O.mpo_to_dense(Ws) - want).max() < 1e-14
    M2 = T.CouplingModel(W1, W2)
    assert len(M2.ids()) == 2 and len(M2) == N


def test_coupling_model_errors():
    os = T.OpStrings()
    os += (1.0, "Sz", 1, "Sz", 2, "Sz", 3)
    H = T.CouplingModel(os, 6)  # fine on the MPS path
    g, sn = T.default_graph_sitenodes(6)
    tree = TT._Tree(6, g, sn, T.find_eccentric_central_node(g, set(sn.values())))
    with pytest.raises(NotImplementedError):
        TT._tree_operator(tree, H)
    with pytest.raises(NotImplementedError):
        TT._tree_operator(tree, T.CouplingModel(O.heisenberg_mpo(6)))
    bad = T.OpStrings()
    bad += (1.0, "Sz", 5)
    with pytest.raises(ValueError, match="Out-of-bounds"):
        T.CouplingModel(bad, 4)
    bad = T.OpStrings()
    bad += (1.0, "Sy", 1)
    with pytest.raises(ValueError, match="unknown operator"):
        T.CouplingModel(bad, 4)
