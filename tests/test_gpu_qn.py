"""GPU parity of the QN-conserving (U(1), U(1)xU(1)) path against the QN oracle
(oracle/tn_oracle_qn.py), exact diagonalisation and closed forms.

Bars: FP64 energies to |dE| <= 1e-10 (1e-9 against ED / closed forms, which
the DMRG reaches to its own convergence), block weights to 1e-13, and the QN
sectors of every link (QN -> dimension) exactly as the oracle's."""

import ctypes as C
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle import tn_oracle_qn as Q

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def T():
    import tennetlib_amd as T
    T.default_context()
    return T


def _qn_mps(rng, sq, links):
    A = []
    for k in range(len(sq)):
        ql, qr = links[k], links[k + 1]
        a = rng.standard_normal((len(ql), len(sq[k]), len(qr)))
        for i, s, j in itertools.product(range(len(ql)), range(len(sq[k])), range(len(qr))):
            if Q.qadd(ql[i], sq[k][s]) != qr[j]:
                a[i, s, j] = 0.0
        A.append(a)
    return A


@pytest.mark.parametrize("ortho", ["left", "right"])
def test_qn_replacebond_matches_oracle(T, ortho):
    """Block factorisation through the C ABI: kept weights, truncation error,
    the new link's sectors and the product of the two new sites."""
    hs = Q.hubbard_qns()
    rng = np.random.default_rng(5)
    links = [[(0, 0)], [(0, 0), (1, 0), (0, 1), (1, 1)],
             [(1, 0), (0, 1), (1, 1), (1, 1), (2, 1), (1, 2), (2, 0)],
             [(2, 1), (1, 2), (2, 2), (2, 2), (3, 1), (1, 3), (2, 1), (3, 2)],
             [(3, 2), (2, 3), (3, 3), (2, 2), (4, 2)], [(4, 4)]]
    sq = [hs] * 5
    A = _qn_mps(rng, sq, links)
    # the last site must reach the single total QN: restrict its entries accordingly
    env = T.StateEnvs(A, O.hubbard_mpo(5, 1.0, 4.0), qns=sq)
    b = 3
    phi = np.tensordot(A[b - 1], A[b], axes=([2], [0]))
    phi /= np.linalg.norm(phi)
    # cutoff 0 would keep rounding-floor weights whose count depends on the
    # eigensolver; 1e-14 drops them in both
    for maxdim, cutoff in [(100, 1e-14), (9, 1e-12), (4, 1e-3)]:
        _, _, newq, spec = Q.block_factorize(phi, links[b - 1], hs, hs, links[b + 1], maxdim, 1, cutoff, ortho)
        env2 = T.StateEnvs(A, O.hubbard_mpo(5, 1.0, 4.0), qns=sq)
        ctx = env2.ctx
        dphi = T.DeviceArray.from_numpy(ctx, phi)
        terr, nk = C.c_double(), C.c_long()
        eigs = (C.c_double * 256)()
        print(f"qn replacebond {ortho} maxdim={maxdim} cutoff={cutoff}", flush=True)
        ctx.check(ctx.lib.tn_replacebond(env2.h, b, dphi.ptr, 1 if ortho == "left" else 0, maxdim, 1, cutoff, 0.0,
                                         1, C.byref(terr), eigs, 256, C.byref(nk)))
        assert nk.value == len(newq)
        assert np.abs(np.array(eigs[:nk.value]) - spec.eigs).max() < 1e-13
        assert abs(terr.value - spec.truncerr) < 1e-13
        assert env2.sectors(b) == {q: n for q, n in spec.sectors.items()}
        got = env2.psi
        prod = np.tensordot(got[b - 1], got[b], axes=([2], [0]))
        want = np.tensordot(*Q.block_factorize(phi, links[b - 1], hs, hs, links[b + 1], maxdim, 1, cutoff,
                                               ortho)[:2], axes=([2], [0]))
        assert np.abs(prod - want).max() < 1e-12
        # the new sites respect the QNs (zeros outside the allowed blocks are exact)
        lq = [env2.link_qns(k) for k in range(6)]
        assert Q.check_flux(got, sq, lq, tol=0.0)
        dphi.free()
        env2.close()
    env.close()


@pytest.mark.parametrize("ortho", ["left", "right"])
def test_qn_replacebond_block_size_classes(T, ortho):
    """U(1) blocks of 10 to 120 rows.  ortho left: 2, 6 and 6 blocks in the
    size classes up to 32, 64 and 96 rows, enough for each class to take the
    batched Jacobi launch; ortho right: blocks of 60 and 120 rows through the
    dense eigensolver.  Kept weights to 1e-13 of the oracle's (numpy eigh per
    block), the new site orthonormal to 1e-13, the product of the two new
    sites to 1e-12."""
    sizes_l = dict(zip(range(-12, 13, 2), [10, 40, 40, 40, 20, 40, 40, 40, 20, 40, 40, 40, 10]))
    sizes_r = {q: 60 for q in range(-14, 15, 2)}
    ql = [(q,) for q in sorted(sizes_l) for _ in range(sizes_l[q])]
    qr = [(q,) for q in sorted(sizes_r) for _ in range(sizes_r[q])]
    qs = Q.spin_half_qns()
    rng = np.random.default_rng(21)
    mask = _sector_mask(ql, qs, qs, qr)
    phi = rng.standard_normal(mask.shape) * mask
    # graded weights, so that truncation cuts inside blocks
    phi *= np.exp(-0.02 * np.arange(len(ql)))[:, None, None, None]
    phi /= np.linalg.norm(phi)
    N, b, chim = 4, 2, 64
    dims = [1, len(ql), chim, len(qr), 1]
    A = [rng.standard_normal((dims[k], 2, dims[k + 1])) for k in range(N)]
    for maxdim, cutoff in [(10 ** 6, 1e-14), (300, 1e-12), (100, 1e-6)]:
        env = T.StateEnvs(A, O.heisenberg_mpo(N))
        ctx = env.ctx
        sq = np.array([x for _ in range(N) for q in qs for x in q], dtype=np.int32)
        links = [[(0,)], ql, [(0,)] * chim, qr, [(0,)]]
        lq = np.array([x for lk in links for q in lk for x in q], dtype=np.int32)
        ctx.check(ctx.lib.tn_env_set_qn(env.h, 1, sq.ctypes.data_as(C.POINTER(C.c_int)),
                                        lq.ctypes.data_as(C.POINTER(C.c_int))))
        env.nq = 1
        Ab, Ab1, newq, spec = Q.block_factorize(phi, ql, qs, qs, qr, maxdim, 1, cutoff, ortho)
        dphi = T.DeviceArray.from_numpy(ctx, phi)
        terr, nk = C.c_double(), C.c_long()
        eigs = (C.c_double * 4096)()
        print(f"qn replacebond classes {ortho} maxdim={maxdim} cutoff={cutoff}", flush=True)
        ctx.check(ctx.lib.tn_replacebond(env.h, b, dphi.ptr, 1 if ortho == "left" else 0, maxdim, 1, cutoff, 0.0,
                                         1, C.byref(terr), eigs, 4096, C.byref(nk)))
        assert nk.value == len(newq)
        assert np.abs(np.array(eigs[:nk.value]) - spec.eigs).max() < 1e-13
        assert env.sectors(b) == spec.sectors
        if maxdim > 10 ** 5:
            # the union of the block spectra is the dense spectrum of phi
            P = np.linalg.svd(phi.reshape(len(ql) * 2, -1), compute_uv=False) ** 2
            P = P[P > 1e-14]
            assert np.abs(np.array(eigs[:len(P)]) - P).max() < 1e-13
        got = [env.get_site(b), env.get_site(b + 1)]
        if ortho == "left":
            U = got[0].reshape(-1, nk.value)
            assert np.abs(U.T @ U - np.eye(nk.value)).max() < 1e-13
        else:
            U = got[1].reshape(nk.value, -1)
            assert np.abs(U @ U.T - np.eye(nk.value)).max() < 1e-13
        prod = np.tensordot(got[0], got[1], axes=([2], [0]))
        assert np.abs(prod - np.tensordot(Ab, Ab1, axes=([2], [0]))).max() < 1e-12
        dphi.free()
        env.close()


def test_qn_dmrg_heisenberg_matches_oracle_and_ed(T):
    N = 12
    sq = [Q.spin_half_qns()] * N
    ref = Q.QNState(O.neel_mps(N), O.heisenberg_mpo(N), sq)
    sw_ref = Q.dmrg(ref, [16, 32, 64, 64], 1e-12)
    env = T.StateEnvs(O.neel_mps(N), O.heisenberg_mpo(N), qns=sq)
    params = T.make_dmrg_params(maxdim=[16, 32, 64, 64], nsweeps=[1] * 4, cutoff=1e-12)
    print("qn dmrg heisenberg n12", flush=True)
    sw = T.dmrg_(env, params, 2, outputlevel=0)
    assert np.abs(np.array(sw.energy) - np.array(sw_ref.energy)).max() < 1e-10
    assert abs(sw.energy[-1] - (-5.14209063284054)) < 1e-9
    for k in range(N + 1):
        assert env.sectors(k) == ref.sectors(k), k


def test_qn_dmrg_hubbard_closed_form_and_ed(T):
    """U = 0, N = 8, sector (4, 4): free fermions; U = 4, N = 6, sector (3, 3): ED."""
    hs = Q.hubbard_qns()
    N = 8
    env = T.StateEnvs(T.hubbard_occupation_state([1, 2] * (N // 2)), T.hubbard_mpo(N, 1.0, 0.0), qns=[hs] * N)
    print("qn dmrg hubbard n8 u0", flush=True)
    sw = T.dmrg_(env, T.make_dmrg_params(maxdim=[16, 64, 256, 256], nsweeps=[1] * 4, cutoff=1e-13), 2,
                 outputlevel=0)
    assert abs(sw.energy[-1] - Q.hubbard_free_fermion_energy(N, 4, 4)) < 1e-9
    assert env.link_qns(N) == [(4, 4)]
    N = 6
    env = T.StateEnvs(T.hubbard_occupation_state([1, 2] * (N // 2)), T.hubbard_mpo(N, 1.0, 4.0), qns=[hs] * N)
    sw = T.dmrg_(env, T.make_dmrg_params(maxdim=[16, 64, 64], nsweeps=[1] * 3, cutoff=1e-13), 2, outputlevel=0)
    H = O.mpo_to_dense(O.hubbard_mpo(N, 1.0, 4.0))
    qs = [tuple(np.sum([hs[x] for x in idx], axis=0)) for idx in itertools.product(range(4), repeat=N)]
    sel = [i for i, q in enumerate(qs) if q == (3, 3)]
    e_ed = np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0]
    assert abs(sw.energy[-1] - e_ed) < 1e-9


def _golden_run(T, name):
    g = json.loads((GOLDEN / f"{name}.json").read_text())
    N = g["N"]
    hs = Q.hubbard_qns()
    env = T.StateEnvs(T.hubbard_occupation_state([1, 2] * (N // 2)), T.hubbard_mpo(N, g["t"], g["U"]),
                      qns=[hs] * N)
    params = T.make_dmrg_params(maxdim=g["maxdim"], nsweeps=[1] * len(g["maxdim"]), cutoff=g["cutoff"])
    print(f"qn dmrg {name}", flush=True)
    sw = T.dmrg_(env, params, 2, outputlevel=0)
    dE = np.abs(np.array(sw.energy) - np.array(g["energy"])) / N
    print(f"{name} |dE/N|", dE, "maxchi", sw.maxchi, g["maxchi"], flush=True)
    want = [{tuple(int(x) for x in q.split(",")): n for q, n in g["sectors"][k].items()} for k in range(N + 1)]
    got = [env.sectors(k) for k in range(N + 1)]
    return g, sw, dE, got, want


def test_qn_dmrg_hubbard_n16_matches_golden(T):
    """Hubbard N = 16, U = 4, (8, 8), maxdim ramp to 512 with the bond
    dimensions set by cutoff 1e-9: FP64 energies to 1e-10 per site, bond
    dimensions and the QN sectors of every link exactly as the oracle's."""
    g, sw, dE, got, want = _golden_run(T, "hubbard_qn_n16")
    assert dE.max() <= 1e-10
    assert sw.maxchi == g["maxchi"]
    assert got == want


def test_qn_dmrg_hubbard_n32_matches_golden(T):
    """Hubbard N = 32, U = 4, (16, 16), maxdim ramp 32, 64, 128, 128, the
    bond dimensions set by maxdim.  The oracle itself moves by up to 9e-8 per
    site on the ramp sweeps and regroups the sectors of several links under
    2e-16 perturbations of its density matrices (the maxdim cut splits
    multiplets; tests/test_oracle_qn.py::test_hubbard_n32_sensitivity), and
    by 7e-7 under 2e-14 perturbations (the backward error of an eigensolver
    of another kind, such as the Jacobi of small blocks here), so: first
    sweep to 1e-10 per site, ramp sweeps to 1e-5, the last sweep to 5e-9;
    maxchi exact; every link's sectors allowed and its dimension within 4 of
    the oracle's."""
    g, sw, dE, got, want = _golden_run(T, "hubbard_qn_n32")
    assert dE[0] <= 1e-10
    assert dE.max() <= 1e-5
    assert dE[-1] <= 5e-9
    assert sw.maxchi == g["maxchi"]
    N = g["N"]
    for k in range(N + 1):
        assert abs(sum(got[k].values()) - sum(want[k].values())) <= 4, k
        # (N_up, N_dn) of sites 1..k: each in 0..k and reachable from the total (N/2, N/2)
        for (nu, nd) in got[k]:
            assert 0 <= nu <= k and 0 <= nd <= k and nu >= k - N // 2 and nd >= k - N // 2, (k, nu, nd)


def _sector_mask(ql, qs1, qs2, qr):
    m = np.zeros((len(ql), len(qs1), len(qs2), len(qr)), dtype=bool)
    for i, s, t, j in itertools.product(range(len(ql)), range(len(qs1)), range(len(qs2)), range(len(qr))):
        m[i, s, t, j] = Q.qadd(Q.qadd(ql[i], qs1[s]), qs2[t]) == qr[j]
    return m


@pytest.mark.parametrize("model", ["hubbard", "heisenberg"])
def test_qn_blocksparse_product_matches_dense(T, model):
    """The QN product (block-sparse GEMMs over the link sectors, uneven blocks)
    against the dense product of the same MPS and MPO without QNs, at every
    position: FP64 to 1e-12, FP32 to 2e-6 and BF16x3 (the FP32 fragments
    split into bf16 hi/lo parts in the block product) to 1e-4 of max|H phi|;
    entries outside the sectors exactly zero."""
    if model == "hubbard":
        N, sq, Ws = 12, [Q.hubbard_qns()] * 12, T.hubbard_mpo(12, 1.0, 4.0)
        psi0 = T.hubbard_occupation_state([1, 2] * 6)
    else:
        N, sq, Ws = 14, [Q.spin_half_qns()] * 14, O.heisenberg_mpo(14)
        psi0 = O.neel_mps(14)
    env = T.StateEnvs(psi0, Ws, qns=sq)
    T.dmrg_(env, T.make_dmrg_params(maxdim=[16, 64, 128], nsweeps=[1] * 3, cutoff=1e-10), 2, outputlevel=0)
    psi = env.psi
    dense = T.StateEnvs(psi, Ws)
    rng = np.random.default_rng(11)
    for pos in list(range(1, N)) + [N // 2]:
        print(f"qn product {model} pos={pos}", flush=True)
        T.position_(env, pos)
        T.position_(dense, pos)
        mask = _sector_mask(env.link_qns(pos - 1), sq[pos - 1], sq[pos], env.link_qns(pos + 1))
        phi = rng.standard_normal(mask.shape) * mask
        ref = T.product(dense, phi, "fp64")
        scale = np.abs(ref).max()
        got = T.product(env, phi, "fp64")
        assert np.abs(got - ref).max() <= 1e-12 * scale, pos
        assert not np.any(got[~mask])
        got32 = T.product(env, phi, "fp32")
        assert np.abs(got32 - ref).max() <= 2e-6 * scale, pos
        assert not np.any(got32[~mask])
        gotb = T.product(env, phi, "bf16x3")
        errb = np.abs(gotb - ref).max()
        assert errb <= 1e-4 * scale, pos
        assert errb > 1e-12 * scale  # really split into bf16 parts
        assert not np.any(gotb[~mask])


@pytest.mark.parametrize("refp", ["fp64", "fp32"])
def test_qn_dmrg_hubbard_bf16x3_energy(T, refp):
    """QN Hubbard N = 16, U = 4, (8, 8), the cutoff-limited ramp of the golden
    run, with the BF16x3 block product and its Lanczos refinement (FP64 or
    FP32 products): the last two sweeps' energies within 1e-9 per site of
    the FP64 form (measured 2e-10 and 4e-10 with FP64 refinement); the
    refinement is one Krylov step more than the FP64 solve, so the
    unconverged first sweeps follow another path (measured 2.1e-5 and 1e-6,
    held to 1e-4); every link's sectors allowed."""
    g = json.loads((GOLDEN / "hubbard_qn_n16.json").read_text())
    N = g["N"]
    hs = Q.hubbard_qns()
    runs = {}
    for prec in ("fp64", "bf16x3"):
        env = T.StateEnvs(T.hubbard_occupation_state([1, 2] * (N // 2)), T.hubbard_mpo(N, g["t"], g["U"]),
                          qns=[hs] * N)
        params = T.make_dmrg_params(maxdim=g["maxdim"], nsweeps=[1] * len(g["maxdim"]), cutoff=g["cutoff"])
        print(f"qn hubbard n16 {prec} refine {refp}", flush=True)
        sw = T.dmrg_(env, params, 2, outputlevel=0, precision=prec, refine_precision=refp)
        runs[prec] = (np.array(sw.energy), [env.sectors(k) for k in range(N + 1)])
    dE = np.abs(runs["bf16x3"][0] - runs["fp64"][0]) / N
    print("qn hubbard n16 bf16x3 |dE/N|", dE, flush=True)
    assert dE[-2:].max() <= 1e-9
    assert dE.max() <= 1e-4
    for k in range(N + 1):
        for (nu, nd) in runs["bf16x3"][1][k]:
            assert 0 <= nu <= k and 0 <= nd <= k and nu >= k - N // 2 and nd >= k - N // 2, (k, nu, nd)


def test_qn_dmrg_hubbard_n64_u0_closed_form(T):
    """Hubbard N = 64, U = 0, (32, 32): QN DMRG ramp to maxdim 1024 against
    the free-fermion ground energy -80.768626322437; the DMRG energy is
    variational (not below it) and within 2e-6 of it (truncation error of
    maxdim 1024, about 1e-8 per bond)."""
    N = 64
    e0 = Q.hubbard_free_fermion_energy(N, N // 2, N // 2)
    assert abs(e0 - (-80.76862632243697)) < 1e-12
    env = T.StateEnvs(T.hubbard_occupation_state([1, 2] * (N // 2)), T.hubbard_mpo(N, 1.0, 0.0),
                      qns=[Q.hubbard_qns()] * N)
    print("qn dmrg hubbard n64 u0", flush=True)
    sw = T.dmrg_(env, T.make_dmrg_params(maxdim=[64, 128, 256, 512, 1024], nsweeps=[1] * 5, cutoff=1e-14), 2,
                 outputlevel=0)
    print("n64 u0 energies", sw.energy, "maxchi", sw.maxchi, flush=True)
    assert sw.energy[-1] > e0 - 1e-10
    assert sw.energy[-1] - e0 < 2e-6
    assert env.link_qns(N) == [(N // 2, N // 2)]

def _sz0_ground_energy(N, J):
    """Lowest eigenvalue of H = sum J_ij S_i . S_j in the S_z = 0 sector
    (sparse Lanczos on the C(N, N/2) basis states)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    states = [x for x in range(2 ** N) if bin(x).count("1") == N // 2]
    index = {x: i for i, x in enumerate(states)}
    rows, cols, vals = [], [], []
    for i, x in enumerate(states):
        diag = 0.0
        for (a, b), v in J.items():
            ba, bb = (x >> (N - 1 - a)) & 1, (x >> (N - 1 - b)) & 1
            diag += v * (0.25 if ba == bb else -0.25)
            if ba != bb:  # S+ S- + S- S+ flips the pair
                y = x ^ (1 << (N - 1 - a)) ^ (1 << (N - 1 - b))
                rows.append(index[y])
                cols.append(i)
                vals.append(0.5 * v)
        rows.append(i)
        cols.append(i)
        vals.append(diag)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    return spl.eigsh(H, k=1, which="SA")[0][0]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-9)])
def test_qn_dmrg_j1j2_cylinder_sz0_reaches_ed(T, prec, tol):
    """J1-J2 Heisenberg (J2 = 0.5) on a 4 x 4 cylinder with conserved S_z: the
    S_z-conserving compressed MPO (w = 14), Neel start, QN two-site DMRG to
    full bond dimension reaches the S_z = 0 exact-diagonalisation energy, the
    last link carries total S_z = 0 only, and every link's sectors are
    reachable."""
    Ly, Lx = 4, 4
    N = Ly * Lx
    J = T.cylinder_couplings(Ly, Lx, 1.0, 0.5)
    e0 = _sz0_ground_energy(N, J)
    Ws = T.j1j2_cylinder_mpo(Ly, Lx, 1.0, 0.5, conserve_sz=True)
    assert max(w.shape[3] for w in Ws[:-1]) == 14
    sq = [T.spin_half_qns()] * N
    env = T.StateEnvs(T.neel_state(N), Ws, qns=sq)
    params = T.make_dmrg_params(maxdim=[32, 64, 128, 256, 256, 256, 256], nsweeps=[1] * 7, cutoff=1e-12)
    print(f"qn j1j2 4x4 {prec}", flush=True)
    sw = T.dmrg_(env, params, 2, outputlevel=0, precision=prec)
    print("qn j1j2 4x4 energies", sw.energy, "ED", e0, "maxchi", sw.maxchi, flush=True)
    assert abs(sw.energy[-1] - e0) / N < tol
    assert set(env.sectors(N)) == {(0,)}
    for k in range(1, N):
        for (m,) in env.sectors(k):
            assert abs(m) <= min(k, N - k) and (m - k) % 2 == 0, (k, m)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 1e-4)])
def test_qn_product_dense_operator_ignores_stale_workspace(T, prec, tol):
    """With a dense compressed operator (the S_z-conserving J1-J2 MPO) the W
    step multiplies every entry of T1, of which the block GEMM writes only the
    allowed blocks: the rest must be zeros.  A dense FP64 product of a NaN
    vector on the same context first fills the shared workspace with NaN; the
    QN product that follows still matches the dense product."""
    Ly, Lx = 4, 4
    N = Ly * Lx
    Ws = T.j1j2_cylinder_mpo(Ly, Lx, 1.0, 0.5, conserve_sz=True)
    sq = [T.spin_half_qns()] * N
    env = T.StateEnvs(T.neel_state(N), Ws, qns=sq)
    T.dmrg_(env, T.make_dmrg_params(maxdim=[16, 32], nsweeps=[1, 1], cutoff=1e-10), 2, outputlevel=0)
    dense = T.StateEnvs(env.psi, Ws)
    pos = N // 2
    T.position_(env, pos)
    T.position_(dense, pos)
    mask = _sector_mask(env.link_qns(pos - 1), sq[pos - 1], sq[pos], env.link_qns(pos + 1))
    phi = np.random.default_rng(4).standard_normal(mask.shape) * mask
    ref = T.product(dense, phi, "fp64")
    poisoned = T.product(dense, np.full(mask.shape, np.nan), "fp64")
    assert np.isnan(poisoned).all()
    got = T.product(env, phi, prec)
    assert np.isfinite(got).all()
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
