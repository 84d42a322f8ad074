"""Host-side mirror of TenNetLib.jl's MPS hot-path API over libtnhip.

Julia's ``f!`` names become ``f_`` (``position_``, ``dmrg_``, ``fullsweep_``,
``tdvpsweep_``, ``update_position_``); arguments, keyword names, defaults and
error messages follow the reference:

  StateEnvs(psi, H)                      src/mps/state_envs.jl:54-60
  position_/set_nsite_/nsite/product     src/mps/state_envs.jl:341-378
  eig_solver / exp_solver                src/base/solver.jl:23-88
  update_position_                       src/mps/update_site.jl:231-277
  SweepData / fullsweep_                 src/mps/sweep.jl:26-35, 94-196
  DMRGParams / dmrg_ / dmrg / dmrg2      src/mps/dmrg.jl:26-102, 148-297
  TDVPEngine / tdvpsweep_                src/mps/tdvp.jl:17-49, 247-277

An MPS is a list of numpy arrays A_k[a, s, b]; an MPO is a list of numpy
arrays W_k[wl, s', s, wr].  All arithmetic runs on the GPU through the C ABI;
numpy only carries data across the boundary.
"""

from __future__ import annotations

import atexit
import ctypes as C
import math
import time
import weakref
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .couplingmodel import CouplingModel, coupling_mpo, mpo_sum_ids

FLOAT64_THRESHOLD = 1e-15  # src/base/global_variables.jl:10
TYPEMAX_INT = 2 ** 63 - 1
# FP64 Lanczos dimension that refines a BF16x3 eig_solver result
BF16X3_REFINE_KDIM = 2


# ---------------------------------------------------------------------------
# context and device buffers
# ---------------------------------------------------------------------------
class Context:
    """One HIP stream on one device (tn_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.tn_ctx_create(device, C.byref(h)))
        self.h = h
        _live_contexts.add(self)

    def close(self):
        if self.h:
            self.lib.tn_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        L.check(rc, self.h)

    def timers(self, reset: bool = False) -> dict:
        ms = (C.c_double * 4)()
        self.check(self.lib.tn_ctx_timers(self.h, ms, int(reset)))
        return dict(matvec=ms[0], env=ms[1], krylov=ms[2], svd=ms[3])

    def enable_timing(self, on: bool = True):
        self.check(self.lib.tn_ctx_enable_timing(self.h, int(on)))


_default_ctx: Optional[Context] = None

# Live handles, released in dependency order at interpreter exit: device
# buffers and environments before the context (stream) they were made on.
_live_children: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    for obj in list(_live_children):
        try:
            obj.free() if isinstance(obj, DeviceArray) else obj.close()
        except Exception:
            pass
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceArray:
    """A caller-owned device buffer (tn_alloc/tn_free) with a numpy shape."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.tn_alloc(ctx.h, max(self.nbytes, 16), C.byref(p)))
        self.ptr = p
        _live_children.add(self)

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        ctx.check(ctx.lib.tn_memcpy_h2d(ctx.h, d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.tn_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr,
                                                  self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.tn_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# StateEnvs{ProjMPO}
# ---------------------------------------------------------------------------
def _precision_code(p) -> int:
    if isinstance(p, int):
        return p
    return L.PRECISIONS[p]


class StateEnvs:
    """Holds the MPS ``psi`` and its projected operator ``PH`` on the GPU
    (src/mps/state_envs.jl:18-60).  ``StateEnvs(psi, H)`` copies psi.
    ``StateEnvs(psi, H, Ms, weight=w)`` adds the penalty projectors of
    previously found states (ProjMPO_MPS2, state_envs.jl:83-106), for
    excited states.  ``H`` is an MPO (list of W[wl, s', s, wr]), a list of
    MPOs (ProjMPOSum2, state_envs.jl:108-134) or a CouplingModel
    (ProjCouplingModel, state_envs.jl:136-175); the last two run as one MPO
    over their term ids with the reference's product and noise term."""

    def __init__(self, psi: Sequence[np.ndarray], H, Ms=None, weight: Optional[float] = None,
                 ctx: Optional[Context] = None, dtype=None, qns=None):
        if Ms is not None:
            if weight is None:
                raise TypeError("StateEnvs(psi, H, Ms; weight): keyword `weight` is required")
            if weight <= 0.0:
                raise ValueError(f"`weight` parameter should be > 0.0 (value passed was `weight={weight}`)")
        self.ctx = ctx or default_context()
        N = len(psi)
        d = [int(a.shape[1]) for a in psi]
        self.model = H
        noise_skip = None
        if isinstance(H, CouplingModel) or (len(H) and not isinstance(H[0], np.ndarray)):
            # ProjCouplingModel (state_envs.jl:136-175) or ProjMPOSum2 for a
            # list of MPOs (state_envs.jl:108-134): one MPO over the term ids
            if isinstance(H, CouplingModel):
                if len(H) != N or (H.mpos and H.d != d) or (not H.mpos and any(x != 2 for x in d)):
                    raise ValueError("StateEnvs: the CouplingModel and the MPS have different sites")
                ids = H.ids()
            else:
                for Hk in H:
                    if len(Hk) != N:
                        raise ValueError("StateEnvs: MPS and MPO lengths differ")
                ids = mpo_sum_ids(H)
            H, skip_l, skip_r = coupling_mpo(ids, N, d)
            if any(skip_l) or any(skip_r):
                noise_skip = (skip_l, skip_r)
        if len(H) != N:
            raise ValueError("StateEnvs: MPS and MPO lengths differ")
        for k, (a, W) in enumerate(zip(psi, H)):
            if W.shape[1] != d[k] or W.shape[2] != d[k]:
                raise ValueError(f"StateEnvs: site indices of MPO and MPS differ at site {k + 1}")
        w = [int(H[0].shape[0])] + [int(W.shape[3]) for W in H]
        if dtype is None:
            cx = any(np.iscomplexobj(a) for a in psi)
            cx = cx or any(np.iscomplexobj(a) for M in (Ms or []) for a in M)
            dtype = np.complex128 if cx else np.float64
        self.dtype = np.dtype(dtype)
        dt = L.TN_C128 if self.dtype == np.complex128 else L.TN_F64
        h = C.c_void_p()
        di = (C.c_int * N)(*d)
        wi = (C.c_int * (N + 1))(*w)
        self.ctx.check(self.ctx.lib.tn_env_create(self.ctx.h, N, di, wi, dt, C.byref(h)))
        self.h = h
        _live_children.add(self)
        self.N = N
        self.d = d
        self.w = w
        for k, W in enumerate(H, start=1):
            Wc = np.ascontiguousarray(W, dtype=np.float64)
            self.ctx.check(self.ctx.lib.tn_env_set_mpo(self.h, k, Wc.ctypes.data_as(L._PD)))
        for k, A in enumerate(psi, start=1):
            self.set_site(k, A)
        self.H = [np.array(W) for W in H]
        # ProjMPO (one MPO, no penalty projectors): the only kind krylov_extend! takes
        self.plain_mpo = Ms is None and not (isinstance(self.model, CouplingModel)
                                             or (len(self.model) and not isinstance(self.model[0], np.ndarray)))
        if noise_skip is not None:
            sl = (C.c_int * (N + 1))(*noise_skip[0])
            sr = (C.c_int * (N + 1))(*noise_skip[1])
            self.ctx.check(self.ctx.lib.tn_env_set_noise_channels(self.h, sl, sr))
        self.Ms = None
        self.nq = 0
        if qns is not None:
            self._set_qns(psi, qns)
        if Ms is not None:
            self.Ms = [[np.array(a) for a in M] for M in Ms]
            self.ctx.check(self.ctx.lib.tn_env_set_projectors(self.h, len(Ms), float(weight)))
            for j, M in enumerate(Ms):
                if len(M) != N:
                    raise ValueError("StateEnvs: projector MPS length differs from psi")
                for k, A in enumerate(M, start=1):
                    if A.shape[1] != d[k - 1]:
                        raise ValueError(f"StateEnvs: site index of projector {j + 1} differs at site {k}")
                    A = np.ascontiguousarray(A, dtype=self.dtype)
                    self.ctx.check(self.ctx.lib.tn_env_set_projector_site(
                        self.h, j, k, A.ctypes.data_as(C.c_void_p), A.shape[0], A.shape[2]))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.lib.tn_env_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.N

    def set_site(self, k: int, A: np.ndarray):
        A = np.ascontiguousarray(A, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.tn_env_set_site(self.h, k, A.ctypes.data_as(C.c_void_p), A.shape[0],
                                                    A.shape[2]))

    def site_dims(self, k: int):
        l, r = C.c_long(), C.c_long()
        self.ctx.check(self.ctx.lib.tn_env_get_site_dims(self.h, k, C.byref(l), C.byref(r)))
        return l.value, r.value

    def get_site(self, k: int) -> np.ndarray:
        l, r = self.site_dims(k)
        out = np.empty((l, self.d[k - 1], r), dtype=self.dtype)
        self.ctx.check(self.ctx.lib.tn_env_get_site(self.h, k, out.ctypes.data_as(C.c_void_p)))
        return out

    @property
    def psi(self) -> List[np.ndarray]:
        return [self.get_site(k) for k in range(1, self.N + 1)]

    @property
    def orthocenter(self) -> int:
        k = C.c_int()
        self.ctx.check(self.ctx.lib.tn_env_orthocenter(self.h, C.byref(k)))
        return k.value

    def set_orthocenter(self, k: int):
        self.ctx.check(self.ctx.lib.tn_env_set_orthocenter(self.h, k))

    def _set_qns(self, psi, qns):
        """QN conservation (ConserveQNs, test/test_MPS_DMRG.jl:10,29): qns[k]
        lists the QN (tuple of 1 or 2 ints) of every basis state of site k.
        Link QNs are read off psi, which must be block-sparse (every link
        element a definite QN) and right-orthonormal from site 1 (product
        states are)."""
        N = self.N
        if len(qns) != N or any(len(q) != self.d[k] for k, q in enumerate(qns)):
            raise ValueError("StateEnvs: qns must list one QN per basis state of every site")
        nq = len(qns[0][0])
        if nq not in (1, 2) or any(len(x) != nq for q in qns for x in q):
            raise ValueError("StateEnvs: QNs must all have 1 or 2 charges")
        links = [[(0,) * nq]]
        for k, a in enumerate(psi):
            a = np.asarray(a)
            qr = [None] * a.shape[2]
            for i, s_, j in zip(*np.nonzero(a)):
                q = tuple(x + y for x, y in zip(links[-1][i], qns[k][s_]))
                if qr[j] is None:
                    qr[j] = q
                elif qr[j] != q:
                    raise ValueError(f"StateEnvs: link {k + 1} element {j} has no definite QN")
            links.append([q if q is not None else (0,) * nq for q in qr])
        sq = np.array([x for q in qns for x in q], dtype=np.int32).reshape(-1)
        lq = np.array([x for q in links for x in q], dtype=np.int32).reshape(-1)
        self.ctx.check(self.ctx.lib.tn_env_set_qn(self.h, nq, sq.ctypes.data_as(L._PI), lq.ctypes.data_as(L._PI)))
        self.nq = nq
        if all(np.asarray(a).shape[0] == 1 and np.asarray(a).shape[2] == 1 for a in psi):
            self.set_orthocenter(1)  # a product state is orthonormal about any site

    def link_qns(self, k: int):
        """QNs of the elements of link k (0..N) of a QN-conserving environment."""
        n = C.c_long()
        self.ctx.check(self.ctx.lib.tn_env_get_link_qn(self.h, k, None, 0, C.byref(n)))
        buf = (C.c_int * max(1, n.value * self.nq))()
        self.ctx.check(self.ctx.lib.tn_env_get_link_qn(self.h, k, buf, n.value * self.nq, C.byref(n)))
        return [tuple(buf[a * self.nq:(a + 1) * self.nq]) for a in range(n.value)]

    def sectors(self, k: int) -> dict:
        """{QN: dimension} of link k."""
        out = {}
        for q in self.link_qns(k):
            out[q] = out.get(q, 0) + 1
        return out

    def phi_dims(self):
        dims = (C.c_long * 4)()
        n = C.c_long()
        self.ctx.check(self.ctx.lib.tn_env_phi_dims(self.h, dims, C.byref(n)))
        return tuple(dims), n.value

    def copy_state(self):
        return [a.copy() for a in self.psi]


def getpsi(sysenv: StateEnvs) -> List[np.ndarray]:
    """getpsi (state_envs.jl:36)."""
    return sysenv.psi


def nsite(sysenv: StateEnvs) -> int:
    v = C.c_int()
    sysenv.ctx.check(sysenv.ctx.lib.tn_env_nsite(sysenv.h, C.byref(v)))
    return v.value


def set_nsite_(sysenv: StateEnvs, n: int) -> StateEnvs:
    sysenv.ctx.check(sysenv.ctx.lib.tn_env_set_nsite(sysenv.h, n))
    return sysenv


def position_(sysenv: StateEnvs, pos: int) -> StateEnvs:
    sysenv.ctx.check(sysenv.ctx.lib.tn_env_position(sysenv.h, pos))
    return sysenv


def orthogonalize_(sysenv: StateEnvs, site: int) -> StateEnvs:
    sysenv.ctx.check(sysenv.ctx.lib.tn_env_orthogonalize(sysenv.h, site))
    return sysenv


def product(sysenv: StateEnvs, v: Union[np.ndarray, DeviceArray], precision="fp64"):
    """product(sysenv, v) = PH(v) at the current position (state_envs.jl:376).
    Accepts a numpy array (returns numpy) or a DeviceArray (returns one)."""
    dims, n = sysenv.phi_dims()
    ctx = sysenv.ctx
    host = isinstance(v, np.ndarray)
    dv = DeviceArray.from_numpy(ctx, np.asarray(v, dtype=sysenv.dtype)) if host else v
    if int(np.prod(dv.shape)) != n:
        raise L.TNHIPError(L.TN_E_SHAPE,
                           "The order of the ProjMPO-ITensor product P*v is not equal to the order of the "
                           f"ITensor v: v has {int(np.prod(dv.shape))} elements, operator domain {dims}")
    out = DeviceArray(ctx, dv.shape, sysenv.dtype)
    ctx.check(ctx.lib.tn_env_matvec(sysenv.h, dv.ptr, out.ptr, _precision_code(precision)))
    if host:
        r = out.numpy()
        out.free()
        dv.free()
        return r
    return out


# ---------------------------------------------------------------------------
# solvers (markers, dispatched inside update_position_)
# ---------------------------------------------------------------------------
def eig_solver(*args, **kwargs):  # src/base/solver.jl:23
    raise TypeError("eig_solver is passed to update_position_/fullsweep_, not called directly")


def exp_solver(*args, **kwargs):  # src/base/solver.jl:66
    raise TypeError("exp_solver is passed to update_position_/fullsweep_, not called directly")


def _params(solver, kwargs) -> L.UpdateParams:
    p = L.UpdateParams()
    sv = 1 if solver is exp_solver else 0
    L.load().tn_update_params_default(C.byref(p), sv)
    ts = kwargs.get("time_step", None)
    if sv == 1:
        if ts is None:
            raise ValueError("`exp_solver()` is not defined with `time_step=nothing` !!")
        ts = complex(ts)
        p.time_step_re, p.time_step_im = ts.real, ts.imag
    p.normalize = int(kwargs.get("normalize", True))
    p.maxdim = int(min(kwargs.get("maxdim", TYPEMAX_INT), TYPEMAX_INT))
    p.mindim = int(kwargs.get("mindim", 1))
    p.cutoff = float(kwargs.get("cutoff", FLOAT64_THRESHOLD))
    p.reverse_step = int(kwargs.get("reverse_step", ts is not None))
    p.krylovdim = int(kwargs.get("solver_krylovdim", p.krylovdim))
    p.maxiter = int(kwargs.get("solver_maxiter", p.maxiter))
    p.tol = float(kwargs.get("solver_tol", p.tol))
    p.check_convergence = int(kwargs.get("solver_check_convergence", False))
    p.precision = _precision_code(kwargs.get("precision", "fp64"))
    # energy of a reduced-precision solve: 1 = Rayleigh quotient of one FP64
    # product (FP32 form), k >= 2 = an FP64 Lanczos of dimension k from the
    # mixed-precision Ritz vector (BF16x3 form, whose ~1e-5 product error
    # otherwise stays in the state)
    ref = kwargs.get("refine_fp64", None)
    if ref is None:
        ref = BF16X3_REFINE_KDIM if p.precision == L.TN_PREC_BF16X3 else 1
    p.refine_fp64 = int(ref)
    # products of that Lanczos refinement: FP64 (default) or FP32
    p.refine_precision = _precision_code(kwargs.get("refine_precision", "fp64"))
    if p.refine_precision not in (L.TN_PREC_FP64, L.TN_PREC_FP32):
        raise ValueError("refine_precision must be 'fp64' or 'fp32'")
    noise = float(kwargs.get("noise", 0.0))
    if noise > 0 and p.reverse_step:
        raise ValueError(f"`updatePosition()` :: `noise={noise}` cannot be greater than zero for "
                         f"`reverse_step=true` !!")
    p.noise = noise
    if kwargs.get("svd_alg", "divide_and_conquer") not in ("divide_and_conquer", "qr_iteration", "recursive"):
        raise ValueError("unknown svd_alg")
    return p


def _update(sysenv: StateEnvs, p, r, pos: int, nsite_: int, ortho: str):
    fn = sysenv.ctx.lib.tn_update_two_site if nsite_ == 2 else sysenv.ctx.lib.tn_update_one_site
    sysenv.ctx.check(fn(sysenv.h, pos, int(ortho == "left"), C.byref(p), C.byref(r)))


def update_position_(sysenv: StateEnvs, solver, pos: int, nsite_: int, ortho: str, **kwargs):
    """update_position!(sysenv, solver, pos, nsite, ortho; kwargs...)
    (src/mps/update_site.jl:231-277).  Returns (energy, truncerr)."""
    if nsite_ not in (1, 2):
        raise ValueError(f"`update_position()` with `nsite={nsite_}` not implemented !!")
    if ortho not in ("left", "right"):
        raise ValueError("ortho must be \"left\" or \"right\"")
    p = _params(solver, kwargs)
    r = L.UpdateResult()
    _update(sysenv, p, r, pos, nsite_, ortho)
    energy = r.energy if kwargs.get("ishermitian", True) else complex(r.energy, r.energy_im)
    return energy, r.truncerr, r


# ---------------------------------------------------------------------------
# SweepData / fullsweep! / dmrg!
# ---------------------------------------------------------------------------
@dataclass
class SweepData:
    """src/mps/sweep.jl:26-35."""
    sweepcount: int = 0
    maxchi: List[int] = field(default_factory=list)
    energy: List[float] = field(default_factory=list)
    entropy: List[float] = field(default_factory=list)
    maxtruncerr: List[float] = field(default_factory=list)
    lasteigs: List[np.ndarray] = field(default_factory=list)


def _entropy(p: np.ndarray) -> float:
    """_entropy (src/mps/measure.jl:4-20)."""
    p = np.asarray(p, dtype=np.float64)
    s = p.sum()
    if abs(s - 1.0) > 100 * FLOAT64_THRESHOLD:
        p = p / s
    q = p[p > 0.0]
    return float(-(q * np.log(q)).sum())


def fullsweep_(sysenv: StateEnvs, solver, nsite_: int, swdata: SweepData, **kwargs):
    """fullsweep!(sysenv, solver, nsite, swdata; kwargs...) (src/mps/sweep.jl:94-196).
    Returns (dE, dS)."""
    if nsite_ not in (1, 2):
        raise ValueError(f"`fullsweep!()` with `nsite={nsite_}` not implemented !!")
    return _sweep(sysenv, solver, swdata, lambda bond: nsite_, kwargs)


def _sweep(sysenv: StateEnvs, solver, swdata: SweepData, pick, kwargs):
    """Left and right half sweeps with the number of sites of every bond's
    update given by pick(bond) (fullsweep!, src/mps/sweep.jl:94-196;
    dynamic_fullsweep!, sweep.jl:257-376)."""
    outputlevel = kwargs.get("outputlevel", 1)
    noise = kwargs.get("noise", 0.0)
    if sysenv.orthocenter != 1:
        orthogonalize_(sysenv, 1)
    N = len(sysenv)
    energy = float("nan")
    maxtruncerr = 0.0
    swdata.sweepcount += 1
    lasteigs: List[np.ndarray] = [None] * (N - 1)
    p = _params(solver, kwargs)
    # The sweep reports the energy of its last update (sweep.jl:137-157).  A
    # Rayleigh-quotient refinement (refine_fp64 = 1) changes only that number,
    # not the state, so with outputlevel <= 1 it runs for the last update only;
    # a Lanczos refinement (>= 2) moves the state and runs at every update.
    # An FP32 Lanczos refinement moves the state at every update; the last
    # update refines in FP64 so that the reported energy is an FP64 Ritz value.
    p_mid = p
    if p.refine_fp64 == 1 and p.precision != L.TN_PREC_FP64 and outputlevel <= 1:
        p_mid = L.UpdateParams.from_buffer_copy(p)
        p_mid.refine_fp64 = 0
    p_last = p
    if p.refine_fp64 >= 2 and p.refine_precision != L.TN_PREC_FP64:
        p_last = L.UpdateParams.from_buffer_copy(p)
        p_last.refine_precision = L.TN_PREC_FP64
    cap = 1 << 16
    eigbuf = (C.c_double * cap)()
    t0 = time.perf_counter()
    r = L.UpdateResult()
    r.eigs = C.cast(eigbuf, C.POINTER(C.c_double))
    r.eigs_cap = cap
    for bond in range(1, N):
        ns = pick(bond)
        _update(sysenv, p_mid, r, bond, ns, "left")
        energy = r.energy
        maxtruncerr = max(r.truncerr, maxtruncerr)
        lasteigs[bond - 1] = np.array(eigbuf[:min(r.nkept, cap)])
        if outputlevel > 1:
            print(f"At left sweep {swdata.sweepcount} bond {bond} => Energy {energy}, Err {r.truncerr:.2g}")
        if ns == 1 and bond == N - 1:
            _update(sysenv, p_mid, r, N, 1, "left")
            energy = r.energy
    swdata.lasteigs = list(lasteigs)
    for bond in range(N - 1, 0, -1):
        ns = pick(bond)
        last = bond == 1 and ns == 2
        _update(sysenv, p_last if last else p_mid, r, bond + 1 if ns == 1 else bond, ns, "right")
        energy = r.energy
        maxtruncerr = max(r.truncerr, maxtruncerr)
        lasteigs[bond - 1] = np.array(eigbuf[:min(r.nkept, cap)])
        if outputlevel > 1:
            print(f"At right sweep {swdata.sweepcount} bond {bond} => Energy {energy}, Err {r.truncerr:.2g}")
        if ns == 1 and bond == 1:
            _update(sysenv, p_last, r, 1, 1, "right")
            energy = r.energy
    sw_time = time.perf_counter() - t0
    swdata.maxchi.append(max(sysenv.site_dims(k)[1] for k in range(1, N + 1)))
    swdata.energy.append(energy)
    mid = lasteigs[N // 2 - 1]
    swdata.entropy.append(_entropy(mid / mid.sum()))
    swdata.maxtruncerr.append(maxtruncerr)
    swdata.lasteigs = lasteigs
    if swdata.sweepcount > 1:
        enerr = swdata.energy[-1] - swdata.energy[-2]
        enterr = swdata.entropy[-1] - swdata.entropy[-2]
    else:
        enerr = enterr = float("nan")
    if outputlevel > 0:
        print(f"At sweep {swdata.sweepcount} => E={swdata.energy[-1]}, S={swdata.entropy[-1]}, "
              f"MaxLinkDim={swdata.maxchi[-1]}, Noise={noise:.2g}")
        print(f"At sweep {swdata.sweepcount} => dE={enerr}, dS={enterr}, MaxErr={maxtruncerr:.2g}, "
              f"Time={sw_time:.3f}")
    return enerr, enterr


def krylov_extend_(sysenv: StateEnvs, **kwargs) -> None:
    """krylov_extend!(sysenv; kwargs...) (src/mps/sweep.jl:378-452): global
    subspace expansion with extension_krylovdim (3) Krylov vectors H^k psi
    from MPO-MPS products (extension_applyH_cutoff 1e-15,
    extension_applyH_maxdim = maxlinkdim + 2) and basis extension with
    extension_cutoff (1e-7).  The StateEnvs must hold one MPO."""
    if not sysenv.plain_mpo:
        raise ValueError("krylov_extend!: the StateEnvs must be created by a single MPO")
    kdim = int(kwargs.get("extension_krylovdim", 3))
    hcut = float(kwargs.get("extension_applyH_cutoff", FLOAT64_THRESHOLD))
    hmax = int(kwargs.get("extension_applyH_maxdim", 0))
    ecut = float(kwargs.get("extension_cutoff", 1e-7))
    t0 = time.perf_counter()
    sysenv.ctx.check(sysenv.ctx.lib.tn_env_krylov_extend(sysenv.h, kdim, hmax, hcut, ecut))
    if kwargs.get("outputlevel", 1) > 0:
        N = len(sysenv)
        print(f"Global Subspace Expansion: KrylovDim={kdim}, applyH Cutoff={hcut:.2g}, applyH MaxDim={hmax or 'maxlinkdim+2'}")
        print(f"Global Subspace Expansion: Cutoff={ecut:.2g}, "
              f"MaxLinkDim={max(sysenv.site_dims(k)[1] for k in range(1, N + 1))}, "
              f"Time={time.perf_counter() - t0:.3f}")


def dynamic_fullsweep_(sysenv: StateEnvs, solver, swdata: SweepData, eigthreshold: float = 1e-12,
                       extendat: Optional[int] = None, **kwargs):
    """dynamic_fullsweep!(sysenv, solver, swdata; eigthreshold, extendat,
    kwargs...) (src/mps/sweep.jl:257-376): two-site updates on bonds whose
    smallest kept weight of the last sweep is >= eigthreshold and whose
    dimension is below maxdim, one-site updates elsewhere; the first sweep
    and every extendat-th one expand the subspace (one-MPO StateEnvs, then a
    one-site sweep) or run a two-site sweep (other projected operators)."""
    maxdim = int(min(kwargs.get("maxdim", TYPEMAX_INT), TYPEMAX_INT))
    if swdata.sweepcount == 0 or (extendat is not None and (swdata.sweepcount + 1) % extendat == 0):
        if sysenv.plain_mpo:
            krylov_extend_(sysenv, **kwargs)
            return fullsweep_(sysenv, solver, 1, swdata, **kwargs)
        return fullsweep_(sysenv, solver, 2, swdata, **kwargs)

    def pick(bond):
        ev = swdata.lasteigs[bond - 1]
        small = len(ev) == 0 or ev[-1] < eigthreshold
        return 1 if (small or sysenv.site_dims(bond)[1] >= maxdim) else 2

    return _sweep(sysenv, solver, swdata, pick, kwargs)


@dataclass
class DMRGParams:
    """src/mps/dmrg.jl:26-33."""
    maxdim: List[int]
    nsweeps: List[int]
    cutoff: List[float]
    noise: List[float]
    noisedecay: List[float]
    disable_noise_after: List[int]


def make_dmrg_params(*, maxdim, nsweeps, cutoff=FLOAT64_THRESHOLD, noise=0.0, noisedecay=1.0,
                     disable_noise_after=-1) -> DMRGParams:
    """DMRGParams(; maxdim, nsweeps, cutoff, noise, noisedecay,
    disable_noise_after) (src/mps/dmrg.jl:72-102)."""
    n = len(nsweeps)
    vec = lambda x, cast: [cast(v) for v in x] if isinstance(x, (list, tuple, np.ndarray)) else [cast(x)] * n
    p = DMRGParams(list(maxdim), list(nsweeps), vec(cutoff, float), vec(noise, float), vec(noisedecay, float),
                   vec(disable_noise_after, int))
    if not (len(p.nsweeps) == len(p.maxdim) == len(p.cutoff) == len(p.noise) == len(p.noisedecay)
            == len(p.disable_noise_after)):
        raise ValueError("`DMRGParams()` :: Size mismatch in input vectors !! \n"
                         "Lengths of `maxdim` and `nsweeps` must be same !!\n"
                         "`cutoff`, `noise`, `noisedecay`, `disable_noise_after` "
                         "can be either scalar or vector having length = `length(maxdim)` !!")
    return p


def dmrg_(sysenv: StateEnvs, params: DMRGParams, nsite_: int = 2, **kwargs) -> SweepData:
    """dmrg!(sysenv, params, nsite; kwargs...) (src/mps/dmrg.jl:148-227)."""
    outputlevel = kwargs.get("outputlevel", 1)
    enerrgoal = kwargs.get("energyErrGoal", None)
    enterrgoal = kwargs.get("entropyErrGoal", None)
    swdata = SweepData()
    kw = {k: v for k, v in kwargs.items() if k not in ("maxdim", "cutoff", "noise")}
    for ii in range(len(params.nsweeps)):
        maxdim = params.maxdim[ii]
        cutoff = params.cutoff[ii]
        noise = params.noise[ii]
        noisedecay = params.noisedecay[ii]
        disable_noise_after = params.disable_noise_after[ii]
        if outputlevel > 0:
            print(f"DMRG level={ii + 1} => maxdim={maxdim}, nsweeps={params.nsweeps[ii]}, cutoff={cutoff:.2E}")
        for jj in range(1, params.nsweeps[ii] + 1):
            enerr, enterr = fullsweep_(sysenv, eig_solver, nsite_, swdata, maxdim=maxdim, cutoff=cutoff,
                                       noise=noise, **kw)
            errGoalMet = False
            if enerrgoal is not None or enterrgoal is not None:
                if enerrgoal is not None and enterrgoal is not None:
                    errGoalMet = abs(enerr) < abs(enerrgoal) and abs(enterr) < abs(enterrgoal)
                elif enerrgoal is not None:
                    errGoalMet = abs(enerr) < abs(enerrgoal)
                # entropy-only goal never sets errGoalMet (src/mps/dmrg.jl:190-191)
            if errGoalMet and abs(noise) < FLOAT64_THRESHOLD:
                break
            if jj == disable_noise_after:
                noise = 0.0
            noise /= noisedecay
            if noise < 100 * FLOAT64_THRESHOLD:
                noise = 0.0
    return swdata


def dmrg(psi0, H, *args, **kwargs):
    """dmrg(psi0, H, params, nsite) (src/mps/dmrg.jl:231-236) or, for excited
    states, dmrg(psi0, H, Ms, params, nsite; weight) (src/mps/dmrg.jl:238-245)
    -> (energy, psi)."""
    if args and isinstance(args[0], DMRGParams):
        Ms, params, rest = None, args[0], args[1:]
    else:
        Ms, params, rest = args[0], args[1], args[2:]
    nsite_ = rest[0] if rest else kwargs.pop("nsite", 2)
    weight = kwargs.pop("weight", -1.0)
    sysenv = StateEnvs(psi0, H) if Ms is None else StateEnvs(psi0, H, Ms, weight=weight)
    sw = dmrg_(sysenv, params, nsite_, **kwargs)
    return sw.energy[-1], sysenv.psi


def dmrg2(psi0, H, *args, **kwargs):
    """dmrg2 (src/mps/dmrg.jl:295-297): dmrg(..., 2)."""
    return dmrg(psi0, H, *args, 2, **kwargs)


def dmrg1(psi0, H, *args, **kwargs):
    """dmrg1 (src/mps/dmrg.jl:314-320): dmrg(..., 1)."""
    return dmrg(psi0, H, *args, 1, **kwargs)


# ---------------------------------------------------------------------------
# TDVP
# ---------------------------------------------------------------------------
class TDVPEngine:
    """src/mps/tdvp.jl:17-49."""

    def __init__(self, psi, H, Ms=None, weight: Optional[float] = None, ctx: Optional[Context] = None):
        """TDVPEngine(psi, H) or TDVPEngine(psi, H, Ms; weight) (tdvp.jl:41-63)."""
        self.sysenv = StateEnvs(psi, H, Ms, weight=weight, ctx=ctx, dtype=np.complex128)
        self.swdata = SweepData()
        self.abstime = 0.0


def tdvpsweep_(engine: TDVPEngine, time_step, nsite_="dynamic", solver=exp_solver, **kwargs):
    """tdvpsweep!(engine, time_step; nsite, solver, kwargs...) (src/mps/tdvp.jl:247-277)."""
    if solver is not exp_solver:
        raise ValueError("`tdvpsweep!()`: `solver` must be `exp_solver` !!")
    if nsite_ == "dynamic":
        dynamic_fullsweep_(engine.sysenv, exp_solver, engine.swdata, time_step=0.5 * time_step,
                           reverse_step=True, **kwargs)
        engine.abstime += abs(time_step)
        return None
    if nsite_ not in (1, 2):
        raise ValueError("`tdvpsweep!()`: `nsite` must be `\"dynamic\"`, `2`, or `1` !!")
    if kwargs.get("extendat", None) is not None:
        raise ValueError("`tdvpsweep!()`: `extendat` must be `nothing` for `nsite == 2` or `nsite == 1`.")
    fullsweep_(engine.sysenv, exp_solver, nsite_, engine.swdata, time_step=0.5 * time_step,
               reverse_step=True, **kwargs)
    engine.abstime += abs(time_step)
