"""ctypes binding of libtnhip.so (the C ABI declared in include/tnhip.h).

The shared library is loaded from the package directory (built in-tree by
``tennetlib.jl_amd/build.py``).  There is no CPU fallback: if the library is
missing or the device is not a gfx950 GPU, calls raise ``TNHIPError``.

``TNHIP_LIB=<path>`` loads another build of the same ABI instead, such as the
host-instrumented ``libtnhip_hardened.so`` (``build.py --hardened``).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ["TNHIP_LIB"]).resolve() if os.environ.get("TNHIP_LIB") else PKG_ROOT / "libtnhip.so"

TN_OK = 0
TN_E_INVALID = -1
TN_E_SHAPE = -2
TN_E_POSITION = -3
TN_E_NOMEM = -4
TN_E_HIP = -5
TN_E_NOTCONV = -6
TN_E_UNSUPPORTED = -7
TN_E_NUMERIC = -8

TN_F64 = 0
TN_C128 = 1
TN_PREC_FP64 = 0
TN_PREC_FP32 = 1
TN_PREC_BF16X3 = 2
PRECISIONS = {"fp64": TN_PREC_FP64, "fp32": TN_PREC_FP32, "bf16x3": TN_PREC_BF16X3}


class TNHIPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[tnhip {code}] {msg}")
        self.code = code


class KrylovInfo(C.Structure):
    _fields_ = [("converged", C.c_int), ("numops", C.c_int), ("numiter", C.c_int),
                ("residual", C.c_double)]


class UpdateParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("time_step_re", C.c_double), ("time_step_im", C.c_double),
                ("normalize", C.c_int), ("maxdim", C.c_long), ("mindim", C.c_long),
                ("cutoff", C.c_double), ("reverse_step", C.c_int), ("krylovdim", C.c_int),
                ("maxiter", C.c_int), ("tol", C.c_double), ("check_convergence", C.c_int),
                ("precision", C.c_int), ("refine_fp64", C.c_int), ("noise", C.c_double),
                ("refine_precision", C.c_int)]


class UpdateResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("energy_im", C.c_double), ("truncerr", C.c_double),
                ("nkept", C.c_long), ("info", KrylovInfo), ("eigs", C.POINTER(C.c_double)),
                ("eigs_cap", C.c_long)]


_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_D = C.c_double
_PL = C.POINTER(C.c_long)
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "tn_ctx_create": [_I, C.POINTER(_P)],
    "tn_ctx_destroy": [_P],
    "tn_last_error": [_P],
    "tn_ctx_synchronize": [_P],
    "tn_alloc": [_P, C.c_size_t, C.POINTER(_P)],
    "tn_free": [_P, _P],
    "tn_memcpy_h2d": [_P, _P, _P, C.c_size_t],
    "tn_memcpy_d2h": [_P, _P, _P, C.c_size_t],
    "tn_ctx_timers": [_P, _PD, _I],
    "tn_ctx_enable_timing": [_P, _I],
    "tn_ctx_enable_timing_mask": [_P, C.c_uint],
    "tn_ctx_phase_stats": [_P, _I, _PD, _PL],
    "tn_ctx_phase_bytes": [_P, _I, _PD],
    "tn_update_params_default": [C.POINTER(UpdateParams), _I],
    "tn_env_create": [_P, _I, _PI, _PI, _I, C.POINTER(_P)],
    "tn_env_destroy": [_P],
    "tn_env_set_mpo": [_P, _I, _PD],
    "tn_env_set_site": [_P, _I, _P, _L, _L],
    "tn_env_get_site_dims": [_P, _I, _PL, _PL],
    "tn_env_get_site": [_P, _I, _P],
    "tn_env_set_orthocenter": [_P, _I],
    "tn_env_orthocenter": [_P, _PI],
    "tn_env_set_nsite": [_P, _I],
    "tn_env_nsite": [_P, _PI],
    "tn_env_position": [_P, _I],
    "tn_env_reset": [_P],
    "tn_env_set_projectors": [_P, _I, _D],
    "tn_env_set_projector_site": [_P, _I, _I, _P, _L, _L],
    "tn_env_set_noise_channels": [_P, _P, _P],
    "tn_env_krylov_extend": [_P, _I, _L, _D, _D],
    "tn_ttn_set_projectors": [_P, _I, _D],
    "tn_ttn_set_projector_node": [_P, _I, _I, C.POINTER(_L), _PD],
    "tn_ttn_create": [_P, _I, C.POINTER(_P)],
    "tn_ttn_destroy": [_P],
    "tn_ttn_set_link": [_P, _I, _I, _I, _I, _I],
    "tn_ttn_set_site": [_P, _I, _I, _I, _I, _PD],
    "tn_ttn_set_coupling": [_P, _I, _I, _PI, _PI, _PI, _PD],
    "tn_ttn_set_node": [_P, _I, _PL, _PD],
    "tn_ttn_get_node": [_P, _I, _PL, _PD],
    "tn_ttn_node_ptr": [_P, _I, C.POINTER(_PD)],
    "tn_ttn_update_env": [_P, _I, _I],
    "tn_ttn_matvec": [_P, _I, _P, _P],
    "tn_ttn_eig": [_P, _I, _I, _I, _D, _PD, C.POINTER(KrylovInfo)],
    "tn_ttn_move": [_P, _I, _I, _L, _L, _D, _PD, _PL],
    "tn_ttn_normalize": [_P, _I, _PD],
    "tn_ttn_node_norm": [_P, _I, _PD],
    "tn_ttn_expand": [_P, _I, _I, _L, _PD, _PD],
    "tn_ttn_set_center": [_P, _I],
    "tn_ttn_get_center": [_P, _PI],
    "tn_ttn_seed": [_P, C.c_ulonglong],
    "tn_ttn_position": [_P, _I, _L, _L, _D, _I, _I],
    "tn_ttn_update_position": [_P, _I, _L, _L, _D, _I, _I, _I, _D, _I, _PD],
    "tn_ttn_expand_random": [_P, _I, _I, _L, _D],
    "tn_ttn_fullsweep": [_P, _PI, _I, _I, _L, _L, _D, _D, _I, _I, _I, _I, _I, _D, _I, _I, _PD, _PL],
    "tn_env_phi_dims": [_P, _PL, _PL],
    "tn_env_matvec": [_P, _P, _P, _I],
    "tn_env_matvec_noalloc": [_P, _P, _P],
    "tn_env_set_qn": [_P, _I, _PI, _PI],
    "tn_env_get_link_qn": [_P, _I, _PI, _L, _PL],
    "tn_env_contract_phi": [_P, _I, _P],
    "tn_env_orthogonalize": [_P, _I],
    "tn_krylov_eig": [_P, _P, _I, _I, _D, _I, _PD, C.POINTER(KrylovInfo)],
    "tn_krylov_expm": [_P, _P, _D, _D, _I, _I, _D, _I, C.POINTER(KrylovInfo)],
    "tn_rayleigh_fp64": [_P, _P, _PD],
    "tn_replacebond": [_P, _I, _P, _I, _L, _L, _D, _D, _I, _PD, _PD, _L, _PL],
    "tn_svd_trunc": [_P, _P, _L, _L, _L, _L, _D, _P, _PD, _P, _PL, _PD],
    "tn_update_two_site": [_P, _I, _I, C.POINTER(UpdateParams), C.POINTER(UpdateResult)],
    "tn_update_one_site": [_P, _I, _I, C.POINTER(UpdateParams), C.POINTER(UpdateResult)],
    "tn_fullsweep": [_P, C.POINTER(UpdateParams), _PD, _PD, _PL, _PD, _L, _PL],
}
_RESTYPE = {"tn_last_error": C.c_char_p, "tn_update_params_default": None}

_lib = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load libtnhip.so (once) and attach prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise TNHIPError(TN_E_UNSUPPORTED,
                         f"{p} not found: build it with `python tennetlib.jl_amd/build.py` "
                         "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    if path is None:
        _lib = lib
    return lib


def check(rc: int, ctx=None):
    if rc != TN_OK:
        lib = load()
        msg = lib.tn_last_error(ctx)
        raise TNHIPError(rc, msg.decode() if msg else "unknown error")
