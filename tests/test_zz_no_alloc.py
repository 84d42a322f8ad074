"""The allocation-free FP64 product (tn_env_matvec_noalloc) against the
allocating product tn_env_matvec(..., TN_PREC_FP64) and the FP64 oracle.

Both entry points run the same GEMM / W-step / GEMM kernels, so their
results must agree bit for bit; against the oracle the FP64 bar of 1e-12
relative holds.  Two calls in a row share the fixed device scratch: the
second, on a smaller input, must not see the first call's leftovers.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import tn_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import tennetlib_amd as T
    T.default_context()
    return T


def rand_mps(N, d, chis, seed, cplx=False):
    rng = np.random.default_rng(seed)
    dims = [1] + list(chis) + [1]
    A = []
    for k in range(N):
        a = rng.standard_normal((dims[k], d, dims[k + 1]))
        if cplx:
            a = a + 1j * rng.standard_normal(a.shape)
        A.append(a / np.sqrt(dims[k] * d))
    return A


def _both(T, env, psi):
    ctx = env.ctx
    dv = T.DeviceArray.from_numpy(ctx, np.ascontiguousarray(psi, dtype=env.dtype))
    o1 = T.DeviceArray(ctx, dv.shape, env.dtype)
    o2 = T.DeviceArray(ctx, dv.shape, env.dtype)
    ctx.check(ctx.lib.tn_env_matvec_noalloc(env.h, dv.ptr, o1.ptr))
    ctx.check(ctx.lib.tn_env_matvec(env.h, dv.ptr, o2.ptr, 0))
    a, b = o1.numpy(), o2.numpy()
    for x in (dv, o1, o2):
        x.free()
    return a, b


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_matches_allocating_product(T, cplx):
    N = 8
    Ws = O.xxz_mpo(N, 1.0, 0.7, np.linspace(-0.4, 0.6, N))
    chis = [2, 4, 16, 136, 96, 16, 4]
    A = rand_mps(N, 2, chis, seed=11, cplx=cplx)
    env = T.StateEnvs(A, Ws)
    ref = O.Envs(Ws, dtype=np.complex128 if cplx else np.float64)
    rng = np.random.default_rng(12)
    # large first, then smaller inputs: the scratch keeps the earlier contents
    for nsite, pos in ((2, 4), (2, 2), (1, 4), (1, 2)):
        print(f"noalloc cplx={cplx} nsite={nsite} pos={pos}", flush=True)
        T.set_nsite_(env, nsite)
        ref.nsite = nsite
        T.position_(env, pos)
        ref.position(A, pos)
        dims, n = env.phi_dims()
        shape = tuple(dims) if nsite == 2 else (dims[0], dims[1], dims[3])
        psi = rng.standard_normal(shape)
        if cplx:
            psi = psi + 1j * rng.standard_normal(shape)
        got, alloc = _both(T, env, psi)
        assert np.array_equal(got, alloc), (nsite, pos, np.abs(got - alloc).max())
        want = ref.product(psi)
        assert np.linalg.norm(got.reshape(want.shape) - want) <= 1e-12 * np.linalg.norm(want)


def test_noalloc_rejects_oversize_and_projectors(T):
    from tennetlib_amd import _lib as L
    N = 4
    Ws = O.xxz_mpo(N, 1.0, 1.0)
    # at position 2: (5 + 5) x 400 x 4 x 400 doubles = 51 MB > 32 MiB
    A = rand_mps(N, 2, [400, 8, 400], seed=3)
    env = T.StateEnvs(A, Ws)
    print("noalloc oversize", flush=True)
    T.position_(env, 2)
    dims, n = env.phi_dims()
    ctx = env.ctx
    dv = T.DeviceArray.from_numpy(ctx, np.ones(dims))
    out = T.DeviceArray(ctx, dims)
    rc = ctx.lib.tn_env_matvec_noalloc(env.h, dv.ptr, out.ptr)
    assert rc == L.TN_E_INVALID
    assert b"scratch" in ctx.lib.tn_last_error(ctx.h)
    # a small environment with a penalty projector
    A2 = rand_mps(N, 2, [2, 4, 2], seed=4)
    env2 = T.StateEnvs(A2, Ws, [A2], weight=2.0)
    print("noalloc projector", flush=True)
    T.position_(env2, 1)
    dims2, _ = env2.phi_dims()
    dv2 = T.DeviceArray.from_numpy(ctx, np.ones(dims2))
    out2 = T.DeviceArray(ctx, dims2)
    assert ctx.lib.tn_env_matvec_noalloc(env2.h, dv2.ptr, out2.ptr) == L.TN_E_UNSUPPORTED
    for x in (dv, out, dv2, out2):
        x.free()
