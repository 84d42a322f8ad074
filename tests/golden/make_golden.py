"""Writes the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

The reference (TenNetLib.jl) ships no fixtures and cannot run here, so these
vectors come from the oracle after it has been pinned by exact
diagonalisation and closed forms (tests/test_oracle.py); they freeze the
config-1 DMRG trajectory, one matvec, the Heisenberg N=100 DMRG ramp to
maxdim 1024, the XXZ N=100 TDVP run and the J1-J2 6 x 12 cylinder ramp to
maxdim 256, so that the GPU path and later changes of the oracle are
compared against fixed numbers.  `python make_golden.py c2 tdvp` rewrites
only the named fixtures.
"""

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from oracle import tn_oracle as O  # noqa: E402


def c1():
    env = O.StateEnvs(O.random_product_mps(20, 2, seed=1), O.heisenberg_mpo(20))
    maxdim = [16, 32, 64, 64, 64]
    p = O.DMRGParams.make(maxdim, [1] * 5, cutoff=1e-10)
    sw = O.dmrg(env, p, krylovdim=3, maxiter=2)
    out = dict(model="heisenberg", N=20, start="random_product_seed1", maxdim=maxdim, cutoff=1e-10,
               krylovdim=3, maxiter=2, tol=1e-14, energy=sw.energy, maxchi=sw.maxchi, entropy=sw.entropy,
               maxtruncerr=sw.maxtruncerr)
    (HERE / "c1_heisenberg_n20.json").write_text(json.dumps(out, indent=1))


def matvec():
    N, Delta = 8, 0.7
    h = np.linspace(-0.5, 0.5, N)
    rng = np.random.default_rng(123)
    dims = [1, 2, 4, 8, 7, 5, 4, 2, 1]
    A = [rng.standard_normal((dims[k], 2, dims[k + 1])) for k in range(N)]
    pos = 4
    env = O.Envs(O.xxz_mpo(N, 1.0, Delta, h))
    env.nsite = 2
    env.position(A, pos)
    psi = rng.standard_normal((dims[pos - 1], 2, 2, dims[pos + 1]))
    Hpsi = env.product(psi)
    np.savez(HERE / "matvec_n8.npz", N=N, Delta=Delta, h=h, pos=pos, psi=psi, Hpsi=Hpsi,
             **{f"A{k}": A[k] for k in range(N)})


def _sz_profile(A):
    """<Sz_i> of a normalised MPS by left transfer matrices (one pass per site)."""
    sz = np.diag([0.5, -0.5])
    N = len(A)
    out = []
    for i in range(N):
        L = np.ones((1, 1), dtype=complex)
        for k in range(N):
            op = sz if k == i else np.eye(2)
            L = np.einsum("ab,asc,st,btd->cd", L, A[k], op, A[k].conj())
        out.append(float(L[0, 0].real))
    return out


def c2():
    """Heisenberg N=100 from Neel, maxdim ramped to 1024 over 8 sweeps,
    cutoff 1e-12, eig_solver defaults (krylovdim 5, maxiter 2)."""
    N = 100
    env = O.StateEnvs(O.neel_mps(N), O.heisenberg_mpo(N))
    maxdim = [64, 128, 256, 512, 1024, 1024, 1024, 1024]
    p = O.DMRGParams.make(maxdim, [1] * len(maxdim), cutoff=1e-12)
    sw = O.dmrg(env, p, krylovdim=5, maxiter=2)
    out = dict(model="heisenberg", N=N, start="neel", maxdim=maxdim, cutoff=1e-12, krylovdim=5, maxiter=2,
               tol=1e-14, energy=sw.energy, maxchi=sw.maxchi, entropy=sw.entropy, maxtruncerr=sw.maxtruncerr,
               linkdims=[int(a.shape[2]) for a in env.psi[:-1]],
               mid_eigs=[float(x) for x in sw.lasteigs[N // 2 - 1]])
    (HERE / "c2_heisenberg_n100.json").write_text(json.dumps(out, indent=1))


def tdvp_xxz_n100():
    """Two-site TDVP of the XXZ chain (N=100, Delta=0.5) from Neel: 40 calls of
    tdvpsweep! with time_step=-0.05im, maxdim 512, cutoff 1e-12."""
    N, Delta, dt, steps = 100, 0.5, 0.05, 40
    Ws = O.xxz_mpo(N, 1.0, Delta)
    env = O.StateEnvs([a.astype(complex) for a in O.neel_mps(N)], Ws)
    sw = O.SweepData()
    sz = {}
    for it in range(1, steps + 1):
        O.tdvpsweep(env, sw, -1j * dt, maxdim=512, cutoff=1e-12)
        if it in (20, 40):
            sz[str(it)] = _sz_profile(env.psi)
    out = dict(model="xxz", N=N, Delta=Delta, start="neel", dt=dt, steps=steps, maxdim=512, cutoff=1e-12,
               energy=sw.energy, maxchi=sw.maxchi, maxtruncerr=sw.maxtruncerr, sz=sz)
    (HERE / "tdvp_xxz_n100.json").write_text(json.dumps(out, indent=1))


def j1j2_6x12():
    """J1-J2 (J2=0.5) Heisenberg on the Ly=6 x Lx=12 cylinder (snake order,
    periodic in y) from Neel, maxdim ramped to 256 over 4 sweeps, cutoff 1e-10."""
    sys.path.insert(0, str(HERE.parent.parent / "tennetlib.jl_amd"))
    from tennetlib_amd import models as M
    Ly, Lx = 6, 12
    Ws = M.j1j2_cylinder_mpo(Ly, Lx, 1.0, 0.5)
    env = O.StateEnvs(M.neel_state(Ly * Lx), Ws)
    maxdim = [32, 64, 128, 256]
    p = O.DMRGParams.make(maxdim, [1] * len(maxdim), cutoff=1e-10)
    sw = O.dmrg(env, p, krylovdim=5, maxiter=2)
    out = dict(model="j1j2_cylinder", Ly=Ly, Lx=Lx, J1=1.0, J2=0.5, start="neel", maxdim=maxdim, cutoff=1e-10,
               krylovdim=5, maxiter=2, energy=sw.energy, maxchi=sw.maxchi, maxtruncerr=sw.maxtruncerr,
               linkdims=[int(a.shape[2]) for a in env.psi[:-1]])
    (HERE / "j1j2_cylinder_6x12.json").write_text(json.dumps(out, indent=1))


def heisenberg_ed_energy(N: int) -> float:
    """Ground energy of the open Heisenberg chain H = sum S_i . S_{i+1} in the
    Sz = 0 sector by Lanczos (scipy eigsh) on the bit-string basis: the
    diagonal is sum_i +-1/4, and every bond with unequal spins flips them with
    amplitude 1/2 (index lookup by binary search in the sorted basis)."""
    import scipy.sparse.linalg as spl
    n = np.arange(1 << N, dtype=np.uint32)
    c = n - ((n >> 1) & 0x55555555)
    c = (c & 0x33333333) + ((c >> 2) & 0x33333333)
    c = (((c + (c >> 4)) & 0x0F0F0F0F) * 0x01010101) >> 24
    states = n[c == N // 2]
    del n, c
    dim = states.size
    diag = np.zeros(dim)
    pairs = []
    for i in range(N - 1):
        bi = (states >> i) & 1
        bj = (states >> (i + 1)) & 1
        same = bi == bj
        diag += np.where(same, 0.25, -0.25)
        src = np.nonzero(~same)[0].astype(np.int32)
        dst = np.searchsorted(states, states[src] ^ np.uint32(3 << i)).astype(np.int32)
        pairs.append((src, dst))
    del states

    def mv(x):
        x = np.asarray(x).reshape(-1)
        y = diag * x
        for src, dst in pairs:
            y[src] += 0.5 * x[dst]
        return y

    op = spl.LinearOperator((dim, dim), matvec=mv, dtype=np.float64)
    v0 = np.random.default_rng(0).standard_normal(dim)
    w = spl.eigsh(op, k=1, which="SA", v0=v0, tol=1e-14, ncv=30)[0]
    return float(w[0])


def ed_n26():
    """Exact ground energy of the open Heisenberg chain N = 26 (Sz = 0, dim
    10,400,600), the chain of the chi = 4096 DMRG sweep; the same routine
    reproduces the N = 12 and N = 20 values pinned in tests/test_oracle.py."""
    e12 = heisenberg_ed_energy(12)
    e20 = heisenberg_ed_energy(20)
    assert abs(e12 - (-5.14209063284054)) < 1e-11 and abs(e20 - (-8.68247333439896)) < 1e-11, (e12, e20)
    e26 = heisenberg_ed_energy(26)
    out = dict(model="heisenberg", N=26, sector="Sz=0", dim=10400600, energy=e26, check_n12=e12, check_n20=e20)
    (HERE / "heisenberg_n26_ed.json").write_text(json.dumps(out, indent=1))


def hubbard_qn(N=32, maxdim=(32, 64, 128, 128), U=4.0, cutoff=1e-12, name="hubbard_qn_n32"):
    """QN-conserving two-site DMRG of the Hubbard chain (t = 1, U = 4, half
    filling, sector (N/2, N/2)) from the occupation product state up, dn,
    up, dn, ...: per-sweep energies, bond dimensions and the QN sectors of
    every link (oracle/tn_oracle_qn.py)."""
    from oracle import tn_oracle_qn as Q
    occ = [1, 2] * (N // 2)
    psi = O.product_mps([np.eye(4)[o] for o in occ])
    st = Q.QNState(psi, O.hubbard_mpo(N, 1.0, U), [Q.hubbard_qns()] * N)
    sw = Q.dmrg(st, list(maxdim), cutoff)
    out = dict(model="hubbard", N=N, t=1.0, U=U, sector=[N // 2, N // 2], start="up,dn,...", maxdim=list(maxdim),
               cutoff=cutoff, krylovdim=5, maxiter=2, energy=sw.energy, maxchi=sw.maxchi, maxtruncerr=sw.maxtruncerr,
               sectors=[{f"{q[0]},{q[1]}": n for q, n in sorted(st.sectors(k).items())} for k in range(N + 1)])
    (HERE / f"{name}.json").write_text(json.dumps(out, indent=1))


def hubbard_qn_n16():
    """Cutoff-limited QN DMRG (N = 16, maxdim ramp to 512, cutoff 1e-9): no
    maxdim cut through a multiplet, so the sectors of every link are fixed by
    the problem, not by rounding."""
    hubbard_qn(N=16, maxdim=(64, 256, 512, 512), cutoff=1e-9, name="hubbard_qn_n16")


FIXTURES = {"c1": c1, "matvec": matvec, "c2": c2, "tdvp": tdvp_xxz_n100, "j1j2": j1j2_6x12, "ed26": ed_n26,
            "hubbard_qn": hubbard_qn, "hubbard_qn_n16": hubbard_qn_n16}

if __name__ == "__main__":
    for name in (sys.argv[1:] or FIXTURES):
        FIXTURES[name]()
        print(f"fixture {name} written", flush=True)
