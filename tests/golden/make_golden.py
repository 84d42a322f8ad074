"""Writes the golden fixtures of this directory from the FP64 oracle.

    python tests/golden/make_golden.py

The reference (TenNetLib.jl) ships no fixtures and cannot run here, so these
vectors come from the oracle after it has been pinned by exact
diagonalisation and closed forms (tests/test_oracle.py); they freeze the
config-1 DMRG trajectory and one matvec so that the GPU path and later
changes of the oracle are compared against fixed numbers.
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


if __name__ == "__main__":
    c1()
    matvec()
    print("fixtures written")
