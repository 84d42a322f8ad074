"""Pins of the CouplingModel / ProjCouplingModel / ProjMPOSum2 oracle
(oracle/tn_oracle_cm.py) -- CPU only.

* The strings of a CouplingModel sum to the Kronecker-built Hamiltonian.
* Products equal P^H H P on the local space of a mixed-canonical complex MPS
  (one and two sites, every position), for OpStrings models with 1-, 2- and
  3-site terms and for sums of MPOs.
* The merged factors of a Heisenberg bond have a closed form: the operator
  Sz Sz + (S+ S- + S- S+)/2 has singular values (1/2, 1/2, 1/2) over the
  orthogonal pairs (Sz, Sz), (S+, S-), (S-, S+), so after the two sweeps the
  left factors are {Sz, S+/sqrt2, S-/sqrt2} / 3^(1/4) and the right ones
  3^(1/4) {Sz, S-/sqrt2, S+/sqrt2} (up to a rotation of the link).
* Noise terms of few-string models have closed forms built from the dense
  block isometries: a string not yet started contributes nothing, a string
  that acts on the kept side contributes |coef|^(k/n)-scaled products of its
  operators, strings completed before the bond share one block (their cross
  terms appear).
* CouplingModel(mpo) gives the ProjMPO noise term (ITensorMPS noiseterm).
"""

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle import tn_oracle_cm as CM
from tests.test_oracle import _complex_mps, _left_iso, _mixed_canonical, _right_iso, xxz_dense

OPS = {"Id": np.eye(2), "Sz": np.diag([0.5, -0.5]), "S+": np.array([[0.0, 1.0], [0.0, 0.0]]),
       "S-": np.array([[0.0, 0.0], [1.0, 0.0]]), "Sx": np.array([[0.0, 0.5], [0.5, 0.0]])}


def kron_term(N, coef, ops):
    mats = [np.eye(2) for _ in range(N)]
    for name, i in ops:
        mats[i - 1] = mats[i - 1] @ OPS[name]
    out = np.ones((1, 1))
    for m in mats:
        out = np.kron(out, m)
    return coef * out


def mixed_model_terms(N, seed):
    rng = np.random.default_rng(seed)
    terms = []
    for j in range(1, N):
        terms += [(1.0, [("Sz", j), ("Sz", j + 1)]), (0.5, [("S+", j), ("S-", j + 1)]),
                  (0.5, [("S-", j), ("S+", j + 1)])]
    for j in range(1, N - 1):
        terms.append((-0.35, [("Sz", j), ("Sz", j + 2)]))
    terms.append((0.6, [("Sx", 1), ("Sz", 3), ("Sx", N)]))
    terms.append((-0.4, [("S+", 2), ("S-", 2)]))
    terms.append((0.25, [("Id", 3), ("Sx", 4)]))
    for j in range(1, N + 1):
        terms.append((float(rng.uniform(-1, 1)), [("Sz", j)]))
    return terms


def test_coupling_model_strings_sum_to_the_hamiltonian():
    N = 6
    terms = mixed_model_terms(N, 1)
    Hd = sum(kron_term(N, c, ops) for c, ops in terms)
    for merge in (True, False):
        M = CM.CMModel.from_terms(terms, N, merge=merge)
        assert np.abs(M.dense() - Hd).max() < 1e-14
    M = CM.CMModel.from_terms(terms, N)
    sup = sorted(p for p, _ in M.ids)
    # one id per support: NN bonds, NNN bonds, the 3-site string, and one
    # single-site id per site (the same-site product and the field merge)
    assert len(sup) == len(set(sup)) == (N - 1) + (N - 2) + 1 + N
    with pytest.raises(ValueError, match="Out-of-bounds"):
        CM.CMModel.from_terms([(1.0, [("Sz", 0)])], N)


def test_heisenberg_bond_factors_closed_form():
    M = CM.CMModel.from_terms([(1.0, [("Sz", 1), ("Sz", 2)]), (0.5, [("S+", 1), ("S-", 2)]),
                               (0.5, [("S-", 1), ("S+", 2)])], 2)
    (pos, (F1, F2)), = M.ids
    assert pos == (1, 2) and F1.shape == (1, 2, 2, 3) and F2.shape == (3, 2, 2, 1)
    q = 3 ** 0.25
    a = [OPS["Sz"] / q, OPS["S+"] / (np.sqrt(2) * q), OPS["S-"] / (np.sqrt(2) * q)]
    b = [q * OPS["Sz"], q * OPS["S-"] / np.sqrt(2), q * OPS["S+"] / np.sqrt(2)]
    gram = lambda ops: sum(np.outer(o.reshape(-1), o.reshape(-1)) for o in ops)  # noqa: E731
    G1 = sum(np.outer(F1[0, :, :, t].reshape(-1), F1[0, :, :, t].reshape(-1)) for t in range(3))
    G2 = sum(np.outer(F2[t, :, :, 0].reshape(-1), F2[t, :, :, 0].reshape(-1)) for t in range(3))
    assert np.abs(G1 - gram(a)).max() < 1e-15
    assert np.abs(G2 - gram(b)).max() < 1e-15


def _env_at(model_or_mpos, A, p, nsite):
    if isinstance(model_or_mpos, CM.CMModel):
        env = CM.ProjCM(model_or_mpos)
    else:
        env = CM.ProjSum(model_or_mpos, dtype=np.complex128)
    env.nsite = nsite
    env.position(A, p)
    return env


@pytest.mark.parametrize("nsite,p", [(2, 1), (2, 2), (2, 4), (2, 6), (1, 1), (1, 3), (1, 7)])
def test_coupling_model_product_equals_projected_hamiltonian(nsite, p):
    N, d = 7, 2
    terms = mixed_model_terms(N, 2)
    model = CM.CMModel.from_terms(terms, N)
    Hd = sum(kron_term(N, c, ops) for c, ops in terms)
    A = _mixed_canonical(_complex_mps(N, d, [2, 4, 6, 6, 4, 2], seed=70 + p), p, nsite)
    env = _env_at(model, A, p, nsite)
    Lm, Rm = _left_iso(A, p), _right_iso(A, p + nsite - 1)
    P = np.kron(np.kron(Lm, np.eye(d ** nsite)), Rm.T)
    Heff = P.conj().T @ Hd @ P
    rng = np.random.default_rng(p)
    shape = (Lm.shape[1],) + (d,) * nsite + (Rm.shape[0],)
    v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = env.product(v).reshape(-1)
    assert np.abs(got - Heff @ v.reshape(-1)).max() < 1e-12 * np.abs(Heff).max()


@pytest.mark.parametrize("p", [1, 3, 5])
def test_mpo_sum_product_equals_projected_hamiltonian(p):
    N, d = 6, 2
    h = [0.2, -0.5, 0.9, 0.1, -0.3, 0.7]
    W1 = O.xxz_mpo(N, 1.0, 0.6)
    W2 = O.xxz_mpo(N, 0.0, 0.0, h)
    A = _mixed_canonical(_complex_mps(N, d, [2, 4, 5, 4, 2], seed=80 + p), p, 2)
    env = _env_at([W1, W2], A, p, 2)
    Lm, Rm = _left_iso(A, p), _right_iso(A, p + 1)
    P = np.kron(np.kron(Lm, np.eye(d * d)), Rm.T)
    Heff = P.conj().T @ xxz_dense(N, 1.0, 0.7 * 0 + 0.6, h) @ P
    v = np.random.default_rng(p).standard_normal((Lm.shape[1], d, d, Rm.shape[0])) + 0j
    assert np.abs(env.product(v).reshape(-1) - Heff @ v.reshape(-1)).max() < 1e-12 * np.abs(Heff).max()
    for ortho in ("left", "right"):
        want = sum(O.noiseterm(_env_at([W], A, p, 2).parts[0], v, ortho) for W in (W1, W2))
        assert np.abs(env.noiseterm(v, ortho) - want).max() < 1e-13


def _block_op(A, p, ops, side):
    """Operator string restricted to the left block (sites < p) or the right
    block (sites > q) as a matrix on the block's bond space, from the dense
    isometries: left  Lm^H O Lm  [a', a];  right  conj(Rm) O Rm^T [b', b]."""
    N = len(A)
    if side == "left":
        Lm = _left_iso(A, p)
        n = p - 1
        mats = [np.eye(2) for _ in range(n)]
        for name, i in ops:
            mats[i - 1] = mats[i - 1] @ OPS[name]
        Om = np.ones((1, 1))
        for m in mats:
            Om = np.kron(Om, m)
        return Lm.conj().T @ Om @ Lm
    Rm = _right_iso(A, p)  # sites p+1..N
    mats = [np.eye(2) for _ in range(N - p)]
    for name, i in ops:
        mats[i - p - 1] = mats[i - p - 1] @ OPS[name]
    Om = np.ones((1, 1))
    for m in mats:
        Om = np.kron(Om, m)
    return Rm.conj() @ Om @ Rm.T


def test_noise_term_closed_forms():
    """N = 6, strings c1 Sx_1 Sz_2, c2 S+_1 S-_4 and c3 Sz_2 (coefficients
    chosen so that sign and |c|^(1/2) scaling matter), bond p = 3 (sites 3, 4):
    left: all started left of the bond; the first and third are completed and
    share the completed block (with their cross terms), the second is open and
    passes site 3:
      X_done = (Lb^H (c1 O1 + c3 O3) Lb) phi,   X_2 = s2 sqrt|c2| (Lb^H S+_1 Lb) phi;
    c4 Sz_4 Sz_5 starts right of the bond and stays out of the left term;
    right: the second string ends at site 4, the fourth reaches it from the
    right, the others never do:
      X_2 = sqrt|c2| (S- on site 4) phi,
      X_4 = s4 |c4| (Sz on site 4) (conj(Rb) Sz_5 Rb^T) phi."""
    N, d, p = 6, 2, 3
    c1, c2, c3, c4 = -0.8, -1.7, 0.9, -0.6
    terms = [(c1, [("Sx", 1), ("Sz", 2)]), (c2, [("S+", 1), ("S-", 4)]), (c3, [("Sz", 2)]),
             (c4, [("Sz", 4), ("Sz", 5)])]
    model = CM.CMModel.from_terms(terms, N)
    A = _mixed_canonical(_complex_mps(N, d, [2, 4, 5, 4, 2], seed=91), p, 2)
    env = _env_at(model, A, p, 2)
    rng = np.random.default_rng(5)
    a, c = A[p - 1].shape[0], A[p].shape[2]
    phi = rng.standard_normal((a, d, d, c)) + 1j * rng.standard_normal((a, d, d, c))
    Ldone = _block_op(A, p, [("Sx", 1), ("Sz", 2)], "left") * c1 + _block_op(A, p, [("Sz", 2)], "left") * c3
    L2 = _block_op(A, p, [("S+", 1)], "left") * np.sign(c2) * np.sqrt(abs(c2))
    X1 = np.einsum("xa,astb->xstb", Ldone, phi).reshape(a * d, -1)
    X2 = np.einsum("xa,astb->xstb", L2, phi).reshape(a * d, -1)
    want = X1 @ X1.conj().T + X2 @ X2.conj().T
    assert np.abs(env.noiseterm(phi, "left") - want).max() < 1e-13 * np.abs(want).max()
    Y = np.sqrt(abs(c2)) * np.einsum("ut,astb->asub", OPS["S-"], phi).reshape(a * d, d * c)
    R4 = _block_op(A, p + 1, [("Sz", 5)], "right")                                # [b', b]
    Y4 = np.sign(c4) * abs(c4) * np.einsum("ut,astb,yb->asuy", OPS["Sz"], phi, R4).reshape(a * d, d * c)
    want = Y.T @ Y.conj() + Y4.T @ Y4.conj()
    assert np.abs(env.noiseterm(phi, "right") - want).max() < 1e-13 * np.abs(want).max()
    # a string entirely right of the bond does not enter the left noise term
    far = CM.CMModel.from_terms([(2.0, [("Sz", 5), ("Sz", 6)])], N)
    assert np.abs(_env_at(far, A, p, 2).noiseterm(phi, "left")).max() == 0.0
    # one completed right of the bond: X_done = phi (conj(Rb) O Rb^T)^T
    Rdone = _block_op(A, p + 1, [("Sz", 5), ("Sz", 6)], "right") * 2.0           # [b', b]
    Z = np.einsum("astb,yb->asty", phi, Rdone).reshape(a * d, d * c)
    got = _env_at(far, A, p, 2).noiseterm(phi, "right")
    assert np.abs(got - Z.T @ Z.conj()).max() < 1e-13 * np.abs(got).max()


@pytest.mark.parametrize("p", [1, 2, 4, 5])
def test_coupling_model_of_an_mpo_gives_the_projmpo_noise(p):
    N, d = 6, 2
    Ws = O.xxz_mpo(N, 1.0, 0.7, [0.2, -0.5, 0.9, 0.1, -0.3, 0.7])
    A = _mixed_canonical(_complex_mps(N, d, [2, 4, 5, 4, 2], seed=100 + p), p, 2)
    env = _env_at(CM.CMModel.from_mpos(Ws), A, p, 2)
    ref = O.Envs(Ws, dtype=np.complex128)
    ref.nsite = 2
    ref.position(A, p)
    phi = np.random.default_rng(p).standard_normal((A[p - 1].shape[0], d, d, A[p].shape[2])) + 0.5j
    assert np.abs(env.product(phi) - ref.product(phi)).max() < 1e-12
    for ortho in ("left", "right"):
        assert np.abs(env.noiseterm(phi, ortho) - O.noiseterm(ref, phi, ortho)).max() < 1e-12


def _heisenberg_terms(N, h=None):
    t = []
    for j in range(1, N):
        t += [(1.0, [("Sz", j), ("Sz", j + 1)]), (0.5, [("S+", j), ("S-", j + 1)]),
              (0.5, [("S-", j), ("S+", j + 1)])]
    if h is not None:
        t += [(float(h[j - 1]), [("Sz", j)]) for j in range(1, N + 1)]
    return t


def test_coupling_model_noisy_dmrg_n12_reaches_ed():
    model = CM.CMModel.from_terms(_heisenberg_terms(12), 12)
    env = O.StateEnvs(O.neel_mps(12), model)
    p = O.DMRGParams.make([64], [6], cutoff=1e-14, noise=1e-3, noisedecay=2.0, disable_noise_after=2)
    sw = O.dmrg(env, p)
    assert abs(sw.energy[-1] - (-5.14209063284054)) < 1e-10
