"""The C-ABI library loads and exports every symbol include/tnhip.h declares,
the ctypes prototypes cover exactly that set, and the host-only entry points
behave (no device work: this file runs without a GPU)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import tennetlib_amd as T
from tennetlib_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "tnhip.h"


def declared_symbols():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tn_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    lib = L.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s


def test_ctypes_table_matches_header():
    assert sorted(L.SIGNATURES) == declared_symbols()


def test_update_params_defaults_follow_reference():
    lib = L.load()
    p = L.UpdateParams()
    lib.tn_update_params_default(C.byref(p), 0)  # eig_solver (src/base/solver.jl:27-35)
    assert (p.krylovdim, p.maxiter, p.tol, p.reverse_step) == (5, 2, 1e-14, 0)
    assert (p.normalize, p.mindim, p.cutoff) == (1, 1, 1e-15)  # update_site.jl:236-241
    lib.tn_update_params_default(C.byref(p), 1)  # exp_solver (solver.jl:72-79)
    assert (p.krylovdim, p.maxiter, p.tol, p.reverse_step) == (30, 100, 1e-12, 1)


def test_struct_layouts():
    # C layout of tn_update_result: 3 doubles, long, krylov info (3 ints + double), pointer, long
    assert C.sizeof(L.KrylovInfo) == 24
    assert C.sizeof(L.UpdateResult) == 8 * 3 + 8 + 24 + 8 + 8
    # ..., precision, refine_fp64 (ints at 84, 88), noise (double at 96), refine_precision (int at 104)
    assert C.sizeof(L.UpdateParams) == 112
    assert L.UpdateParams.noise.offset == 96
    assert L.UpdateParams.refine_precision.offset == 104


def test_dmrg_params_validation_matches_reference():
    p = T.make_dmrg_params(maxdim=[16, 32], nsweeps=[1, 2], cutoff=1e-10)
    assert p.cutoff == [1e-10, 1e-10] and p.noise == [0.0, 0.0] and p.disable_noise_after == [-1, -1]
    with pytest.raises(ValueError, match="Size mismatch"):
        T.make_dmrg_params(maxdim=[16, 32], nsweeps=[1])
    with pytest.raises(ValueError, match="Size mismatch"):
        T.make_dmrg_params(maxdim=[16], nsweeps=[1], cutoff=[1e-10, 1e-12])


def test_update_position_rejects_noise_with_reverse_step():
    with pytest.raises(ValueError, match="cannot be greater than zero"):
        T.core._params(T.exp_solver, dict(time_step=-0.1j, noise=1e-3))
    with pytest.raises(ValueError, match="time_step"):
        T.core._params(T.exp_solver, {})


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(L.TNHIPError, match="no CPU fallback"):
        L.load(tmp_path / "libtnhip.so")


def test_tnhip_lib_selects_another_build(tmp_path):
    """TNHIP_LIB points the binding at another build of the same ABI (such as the
    host-instrumented one); a path that does not exist fails loudly on load."""
    import os
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, sys.argv[1]); from tennetlib_amd import _lib as L\n"
            "print(L.LIB_PATH)\n"
            "try:\n    L.load()\nexcept L.TNHIPError as e:\n    print('refused', 'no CPU fallback' in str(e))\n")
    want = tmp_path / "libtnhip_other.so"
    env = dict(os.environ, TNHIP_LIB=str(want))
    out = subprocess.run([sys.executable, "-c", code, str(ROOT / "tennetlib.jl_amd")], env=env,
                         capture_output=True, text=True, check=True).stdout.split("\n")
    assert out[0] == str(want.resolve())
    assert out[1] == "refused True"
