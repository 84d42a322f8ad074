"""Build libtnhip.so in-tree for gfx950 (MI355X).

    python tennetlib.jl_amd/build.py              # libtnhip.so
    python tennetlib.jl_amd/build.py --hardened   # libtnhip_hardened.so

Compiles every HIP source under csrc/ (one hipcc process per source, in
parallel) and links them into one shared library next to this file.
Packed-FP32 instructions are disabled for all device code.

The hardened variant instruments the host code only: UndefinedBehaviorSanitizer
(recovering, so every report is logged and the run continues; reports go to
stderr or to UBSAN_OPTIONS=log_path=...) and libstdc++'s precondition checks
(_GLIBCXX_ASSERTIONS: bounds of std::vector::operator[], front/back of empty
containers, ...), which abort on the first violation. Device code is compiled
exactly as in the product build. Load it with TNHIP_LIB=<path> (see _lib.py).
AddressSanitizer and ThreadSanitizer need their runtime loaded before the
Python interpreter starts and are not offered here.
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
OBJ_HARDENED = HERE / "_build_hardened"
OUT_HARDENED = HERE / "libtnhip_hardened.so"
SOURCES = ["gemm.hip", "gemm_fast.hip", "tensor_ops.hip", "svd.hip", "eig_dc.hip", "qr.hip", "tnhip_api.hip",
           "ttn.hip", "qn.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]
# host-side instrumentation of the hardened build (every -fsanitize= behind
# -Xarch_host, so the device code is untouched)
HARDEN_FLAGS = ["-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize=vptr",
                "-Xarch_host", "-fno-omit-frame-pointer", "-D_GLIBCXX_ASSERTIONS", "-g1"]


def ubsan_runtime(cc: str) -> Path:
    """The shared UBSan runtime of hipcc's clang (linked into the hardened library,
    whose host (Python) was not built with the sanitizer)."""
    r = subprocess.run([cc, "--print-file-name=libclang_rt.ubsan_standalone-x86_64.so"], capture_output=True,
                       text=True)
    p = Path(r.stdout.strip())
    if not p.is_absolute() or not p.exists():
        root = Path(cc).resolve().parent.parent
        hits = sorted(root.glob("lib/llvm/lib/clang/*/lib/linux/libclang_rt.ubsan_standalone-x86_64.so"))
        if not hits:
            raise RuntimeError("UBSan runtime (libclang_rt.ubsan_standalone-x86_64.so) not found")
        p = hits[-1]
    return p


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


def build(verbose: bool = False, hardened: bool = False) -> Path:
    cc = hipcc()
    obj_dir, out = (OBJ_HARDENED, OUT_HARDENED) if hardened else (OBJ, OUT)
    flags = FLAGS + (HARDEN_FLAGS if hardened else [])
    obj_dir.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    objs = [obj_dir / (Path(s).stem + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [cc, *flags, "-c", str(CSRC / s), "-o", str(o)], verbose)
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = out.with_suffix(".so.tmp")
    link = [cc, "--offload-arch=gfx950", "-fPIC", "-shared", *[str(o) for o in objs]]
    if hardened:
        rt = ubsan_runtime(cc)
        link += [str(rt), f"-Wl,-rpath,{rt.parent}"]
    _run([*link, "-o", str(tmp)], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose=True, hardened="--hardened" in sys.argv[1:])
    print(f"built {p}")
