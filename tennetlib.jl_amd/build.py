"""Build libtnhip.so in-tree for gfx950 (MI355X).

    python tennetlib.jl_amd/build.py

Compiles every HIP source under csrc/ (one hipcc process per source, in
parallel) and links them into one shared library next to this file.
Packed-FP32 instructions are disabled for all device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
OUT = HERE / "libtnhip.so"
SOURCES = ["gemm.hip", "gemm_fast.hip", "tensor_ops.hip", "svd.hip", "eig_dc.hip", "qr.hip", "tnhip_api.hip",
           "ttn.hip", "qn.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with exit code {r.returncode}: {' '.join(cmd[-3:])}")


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    OBJ.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    objs = [OBJ / (Path(s).stem + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [cc, *FLAGS, "-c", str(CSRC / s), "-o", str(o)], verbose)
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = OUT.with_suffix(".so.tmp")
    _run([cc, "--offload-arch=gfx950", "-fPIC", "-shared", *[str(o) for o in objs], "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    p = build(verbose=True)
    print(f"built {p}")
