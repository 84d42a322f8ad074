"""CPU pins of the host-side model inputs (MPO builders): each MPO, contracted
to a dense matrix, equals the Hamiltonian built independently from Kronecker
products of the same couplings."""

import numpy as np
import pytest

from oracle import tn_oracle as O
from tennetlib_amd import models as M

SZ = np.diag([0.5, -0.5])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])
SM = SP.T.copy()


def _site(op, i, N):
    out = np.ones((1, 1))
    for k in range(N):
        out = np.kron(out, op if k == i else np.eye(2))
    return out


def _dense_from_couplings(N, J, h=None):
    H = np.zeros((2 ** N, 2 ** N))
    for (i, j), v in J.items():
        H += v * (_site(SZ, i, N) @ _site(SZ, j, N)
                  + 0.5 * (_site(SP, i, N) @ _site(SM, j, N) + _site(SM, i, N) @ _site(SP, j, N)))
    if h is not None:
        for i, v in enumerate(h):
            H += v * _site(SZ, i, N)
    return H


@pytest.mark.parametrize("compress", [False, True])
def test_coupling_mpo_random_graph(compress):
    rng = np.random.default_rng(3)
    N = 8
    J = {}
    for _ in range(12):
        i, j = rng.choice(N, 2, replace=False)
        J[(int(min(i, j)), int(max(i, j)))] = float(rng.normal())
    h = rng.normal(size=N)
    Ws = M.spin_coupling_mpo(N, J, h=h, compress=compress)
    assert np.abs(O.mpo_to_dense(Ws) - _dense_from_couplings(N, J, h)).max() < 1e-12


@pytest.mark.parametrize("snake", [True, False])
def test_j1j2_cylinder_mpo(snake):
    Ly, Lx = 3, 3
    J = M.cylinder_couplings(Ly, Lx, 1.0, 0.5, snake)
    # 9 sites: 9 vertical (periodic), 6 horizontal, 12 diagonal bonds
    assert len(J) == 27
    Hd = _dense_from_couplings(Ly * Lx, J)
    for comp in (False, True):
        Ws = M.j1j2_cylinder_mpo(Ly, Lx, 1.0, 0.5, snake=snake, compress=comp)
        assert np.abs(O.mpo_to_dense(Ws) - Hd).max() < 1e-12


def _channel_flux_ok(Ws):
    """Every bond channel carries one 2 S_z flux: flux(b) = flux(a) + q(s) - q(t)."""
    q = [1, -1]
    fl = {0: 0}
    for W in Ws:
        nxt = {}
        for a, s, t, b in zip(*np.nonzero(np.abs(W) > 0)):
            f = fl[int(a)] + q[s] - q[t]
            if nxt.setdefault(int(b), f) != f:
                return False
        fl = nxt
    return fl == {0: 0}


def test_flux_preserving_compression():
    """The S_z-conserving compression gives the same Hamiltonian and the same
    bond dimensions as the plain one, with every channel of definite flux; the
    plain SVD compression mixes the S+, S- and S_z channels."""
    rng = np.random.default_rng(5)
    N = 9
    J = {}
    for _ in range(14):
        i, j = rng.choice(N, 2, replace=False)
        J[(int(min(i, j)), int(max(i, j)))] = float(rng.normal())
    h = rng.normal(size=N)
    Hd = _dense_from_couplings(N, J, h)
    Wp = M.spin_coupling_mpo(N, J, h=h)
    Wq = M.spin_coupling_mpo(N, J, h=h, conserve_sz=True)
    assert np.abs(O.mpo_to_dense(Wq) - Hd).max() < 1e-12
    assert [w.shape for w in Wq] == [w.shape for w in Wp]
    assert _channel_flux_ok(Wq)
    assert not _channel_flux_ok(Wp)
    assert _channel_flux_ok(M.spin_coupling_mpo(N, J, h=h, compress=False))
    Ly, Lx = 3, 3
    Wc = M.j1j2_cylinder_mpo(Ly, Lx, 1.0, 0.5, conserve_sz=True)
    assert _channel_flux_ok(Wc)
    assert np.abs(O.mpo_to_dense(Wc) - _dense_from_couplings(Ly * Lx, M.cylinder_couplings(Ly, Lx, 1.0, 0.5))
                  ).max() < 1e-12


def test_chain_builder_matches_fsm_heisenberg():
    N = 8
    J = {(i, i + 1): 1.0 for i in range(N - 1)}
    Ws = M.spin_coupling_mpo(N, J)
    assert max(w.shape[3] for w in Ws[:-1]) == 5  # compressed to the FSM bond dimension
    assert np.abs(O.mpo_to_dense(Ws) - O.mpo_to_dense(O.heisenberg_mpo(N))).max() < 1e-12


def _fermion_ops(N):
    """c_{i,s} on N sites (modes ordered 1up, 1dn, 2up, ...) by Jordan-Wigner
    from the site operators; checked to satisfy {c_a, c_b^+} = delta_ab."""
    from tennetlib_amd import models as M
    I, F, cup, cdn = M.hubbard_site_ops()
    ops = {}
    for i in range(N):
        for s, c in (("up", cup), ("dn", cdn)):
            mats = [F] * i + [c] + [I] * (N - i - 1)
            op = mats[0]
            for m in mats[1:]:
                op = np.kron(op, m)
            ops[(i, s)] = op
    return ops


def test_hubbard_mpo_is_second_quantised_hamiltonian():
    """The package's Jordan-Wigner MPO equals -t sum (c+_i c_{i+1} + h.c.) +
    U sum n_up n_dn built from fermion operators whose anticommutators are
    checked; at U = 0 the sector ground energies are free-fermion sums."""
    import itertools
    from oracle import tn_oracle as O
    from tennetlib_amd import models as M
    N, t, U = 4, 0.7, 3.1
    c = _fermion_ops(N)
    keys = list(c)
    dim = 4 ** N
    for a, b in itertools.product(keys, keys):
        ac = c[a] @ c[b].T + c[b].T @ c[a]
        assert np.abs(ac - (np.eye(dim) if a == b else 0)).max() < 1e-12
        assert np.abs(c[a] @ c[b] + c[b] @ c[a]).max() < 1e-12
    H = np.zeros((dim, dim))
    for i in range(N - 1):
        for s in ("up", "dn"):
            hop = c[(i, s)].T @ c[(i + 1, s)]
            H -= t * (hop + hop.T)
    for i in range(N):
        H += U * (c[(i, "up")].T @ c[(i, "up")]) @ (c[(i, "dn")].T @ c[(i, "dn")])
    assert np.abs(O.mpo_to_dense(M.hubbard_mpo(N, t, U)) - H).max() < 1e-12
    # U = 0: ground energy of the (2, 2) sector
    H0 = O.mpo_to_dense(M.hubbard_mpo(N, 1.0, 0.0))
    qs = [tuple(np.sum([M.hubbard_qns()[x] for x in idx], axis=0)) for idx in itertools.product(range(4), repeat=N)]
    sel = [k for k, q in enumerate(qs) if q == (2, 2)]
    e = np.linalg.eigvalsh(H0[np.ix_(sel, sel)])[0]
    lv = sorted(-2.0 * np.cos(np.pi * k / (N + 1)) for k in range(1, N + 1))
    assert abs(e - 2 * (lv[0] + lv[1])) < 1e-12
