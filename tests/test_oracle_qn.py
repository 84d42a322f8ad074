"""Pins of the QN-conserving oracle (oracle/tn_oracle_qn.py): block
factorisation against the dense spectrum, the NDTensors block truncation rule
on multiplets split across sectors, closed forms and exact diagonalisation."""

import itertools
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle import tn_oracle_qn as Q


def _qn_phi(rng, ql, qs1, qs2, qr):
    a, s, t, c = len(ql), len(qs1), len(qs2), len(qr)
    phi = rng.standard_normal((a, s, t, c))
    for i, j, k, m in itertools.product(range(a), range(s), range(t), range(c)):
        if Q.qadd(Q.qadd(ql[i], qs1[j]), qs2[k]) != qr[m]:
            phi[i, j, k, m] = 0.0
    return phi / np.linalg.norm(phi)


@pytest.mark.parametrize("ortho", ["left", "right"])
def test_block_factorize_spectrum_and_reconstruction(ortho):
    rng = np.random.default_rng(3)
    hs = Q.hubbard_qns()
    ql = [(0, 0), (1, 0), (1, 0), (0, 1), (1, 1), (1, 1), (2, 1)]
    qr = [(2, 2), (1, 2), (2, 1), (2, 1), (3, 2), (2, 3), (1, 1)]
    phi = _qn_phi(rng, ql, hs, hs, qr)
    Ab, Ab1, newq, spec = Q.block_factorize(phi, ql, hs, hs, qr, 10 ** 9, 1, 0.0, ortho)
    M = phi.reshape(len(ql) * 4, 4 * len(qr))
    P = np.linalg.svd(M, compute_uv=False) ** 2
    P = P[P > 1e-14]
    E = np.sort(spec.eigs)[::-1]
    # blocks of deficient rank give weights at the rounding floor (kept: > docut = 0)
    assert np.all(E[len(P):] < 1e-14)
    assert np.abs(E[:len(P)] - P).max() < 1e-14
    # untruncated: the two new sites reproduce phi, and the new link is QN-definite
    prod = np.tensordot(Ab, Ab1, axes=([2], [0]))
    assert np.abs(prod - phi).max() < 1e-13
    assert Q.check_flux([Ab], [hs], [ql, newq]) and Q.check_flux([Ab1], [hs], [newq, qr])
    # the orthonormal factor is orthonormal
    if ortho == "left":
        U = Ab.reshape(-1, Ab.shape[2])
        assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < 1e-13
    else:
        V = Ab1.reshape(Ab1.shape[0], -1)
        assert np.abs(V @ V.T - np.eye(V.shape[0])).max() < 1e-13


def test_block_truncation_drops_split_multiplets():
    """NDTensors' block rule (truncate!! then P > docut per block): a maxdim
    that cuts through equal weights held by different sectors drops the whole
    multiplet (docut is bumped 1e-3 P[n] above them), where the dense rule
    keeps exactly maxdim; without a cut docut is 0 and exact zeros go."""
    blocks = {(1,): np.array([0.5, 0.0]), (-1,): np.array([0.25]), (3,): np.array([0.25])}
    kept, terr = Q.block_truncate(blocks, 2, 1, 0.0)
    assert kept == {(1,): 1} and abs(terr - 0.25) < 1e-15   # truncate!! error: the one weight beyond maxdim
    kept, terr = Q.block_truncate(blocks, 3, 1, 0.0)
    assert kept == {(1,): 1, (-1,): 1, (3,): 1} and terr == 0.0   # the exact zero is dropped
    # unequal neighbours at the cut: docut halfway, the dense count
    blocks = {(1,): np.array([0.5]), (-1,): np.array([0.3]), (3,): np.array([0.2])}
    kept, terr = Q.block_truncate(blocks, 2, 1, 0.0)
    assert kept == {(1,): 1, (-1,): 1} and abs(terr - 0.2) < 1e-15
    # relative cutoff: drop the tail while its summed weight stays below 0.25 sum(P)
    kept, terr = Q.block_truncate(blocks, 10, 1, 0.25)
    assert kept == {(1,): 1, (-1,): 1}


def test_qn_dmrg_heisenberg_n12_reaches_ed():
    N = 12
    st = Q.QNState(O.neel_mps(N), O.heisenberg_mpo(N), [Q.spin_half_qns()] * N)
    sw = Q.dmrg(st, [16, 32, 64, 64], 1e-12)
    assert abs(sw.energy[-1] - (-5.14209063284054)) < 1e-9
    assert Q.check_flux(st.psi, st.sq, st.links, tol=1e-13)
    assert all(q[0] == 0 for q in st.links[N])  # total 2Sz = 0 kept
    # centre bond: the Sz=0 ground state populates sectors 2Sz = -2j .. 2j
    assert set(st.sectors(N // 2)) <= {(x,) for x in range(-6, 7, 2)}


def test_qn_dmrg_hubbard_free_fermions_and_ed():
    """U = 0 at half filling, N = 8, sector (4, 4): the free-fermion closed
    form; U = 4, N = 6, sector (3, 3): the dense Hamiltonian restricted to the
    sector."""
    hs = Q.hubbard_qns()
    N = 8
    occ = [1, 2] * (N // 2)  # |up>, |dn> alternating: (N_up, N_dn) = (4, 4)
    psi = O.product_mps([np.eye(4)[o] for o in occ])
    st = Q.QNState(psi, O.hubbard_mpo(N, 1.0, 0.0), [hs] * N)
    sw = Q.dmrg(st, [16, 64, 256, 256], 1e-13)
    e0 = Q.hubbard_free_fermion_energy(N, 4, 4)
    assert abs(sw.energy[-1] - e0) < 1e-9, (sw.energy, e0)
    assert st.links[N] == [(4, 4)]
    N = 6
    occ = [1, 2] * (N // 2)
    psi = O.product_mps([np.eye(4)[o] for o in occ])
    st = Q.QNState(psi, O.hubbard_mpo(N, 1.0, 4.0), [hs] * N)
    sw = Q.dmrg(st, [16, 64, 64], 1e-13)
    H = O.mpo_to_dense(O.hubbard_mpo(N, 1.0, 4.0))
    qs = [sum(np.array(hs[x]) for x in idx) for idx in itertools.product(range(4), repeat=N)]
    sel = [i for i, q in enumerate(qs) if tuple(q) == (3, 3)]
    e_ed = np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0]
    assert abs(sw.energy[-1] - e_ed) < 1e-9, (sw.energy, e_ed)


def test_free_fermion_closed_form_n64():
    # the value pinned for the N = 64, U = 0 half-filled chain
    assert abs(Q.hubbard_free_fermion_energy(64, 32, 32) - (-80.76862632243697)) < 1e-10


@pytest.mark.parametrize("amp", [2e-16, 2e-14])
def test_hubbard_n32_sensitivity(monkeypatch, amp):
    """The QN oracle's own rounding sensitivity on the maxdim-limited Hubbard
    N = 32 ramp (tests/golden/hubbard_qn_n32.json): one run whose every block
    density matrix gets a symmetric perturbation of amp max|rho| before its
    eigendecomposition (2e-16: one FP64 rounding; 2e-14: the backward error
    of a different eigensolver).  The ramp sweeps move by more than 1e-8 (at
    2e-14: 1e-7) per site and the sectors of some links regroup, while
    maxchi, the link dimensions to within 4 and the last sweep to 5e-9 per
    site hold: the bounds the GPU comparison uses."""
    g = json.loads((Path(__file__).resolve().parent / "golden" / "hubbard_qn_n32.json").read_text())
    N = g["N"]
    hs = Q.hubbard_qns()
    rng = np.random.default_rng(0)
    eigh = np.linalg.eigh

    def perturbed(r):
        E = rng.standard_normal(r.shape)
        E = 0.5 * (E + E.T)
        return eigh(r + E * (amp * np.abs(r).max() / max(np.abs(E).max(), 1e-300)))

    monkeypatch.setattr(Q.np.linalg, "eigh", perturbed)
    psi = O.product_mps([np.eye(4)[o] for o in [1, 2] * (N // 2)])
    st = Q.QNState(psi, O.hubbard_mpo(N, g["t"], g["U"]), [hs] * N)
    sw = Q.dmrg(st, g["maxdim"], g["cutoff"])
    monkeypatch.undo()
    dE = np.abs(np.array(sw.energy) - np.array(g["energy"])) / N
    print("perturbed |dE/N|", dE)
    assert dE[0] < 1e-10 and dE[1] > (1e-8 if amp < 1e-15 else 1e-7)
    assert dE.max() < 1e-5 and dE[-1] < 5e-9
    assert sw.maxchi == g["maxchi"]
    want = [{tuple(int(x) for x in q.split(",")): n for q, n in g["sectors"][k].items()} for k in range(N + 1)]
    got = [st.sectors(k) for k in range(N + 1)]
    assert sum(a != b for a, b in zip(got, want)) > 0
    assert max(abs(sum(a.values()) - sum(b.values())) for a, b in zip(got, want)) <= 4
