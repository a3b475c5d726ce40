"""Build the gfx950 HIP library ``warp_amd/_lib/libwarp_amd.so``.

Every ``*.hip`` file in this directory is compiled for gfx950 only with hipcc
and linked into one shared library exposing the C ABI declared in
``include/warp_amd.h``.  Run ``python -m warp_amd.csrc.build`` (or
``__graft_entry__.build()``); the build always recompiles.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_DIR = os.path.join(PKG, "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB_NAME = "libwarp_amd.so"

ARCH = "gfx950"

COMPILE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # FMA contraction everywhere, except inside blocks marked "#pragma clang fp contract(off)"
    "-ffp-contract=fast-honor-pragmas",
    # keep the compiler from emitting packed-FP32 VALU instructions
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build warp_amd")
    return path


def sources() -> list[str]:
    return sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".hip"))


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [hipcc(), *COMPILE_FLAGS, *extra, "-I", HERE, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{res.stderr}")
    return obj


def build(verbose: bool = True, extra_flags: list[str] | None = None) -> str:
    """Compile all HIP sources and link the shared library; return its path."""
    t0 = time.time()
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    extra = list(extra_flags or [])
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), srcs))
    out = os.path.join(LIB_DIR, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"[warp_amd] built {out} from {len(srcs)} sources in {time.time() - t0:.1f}s")
    return out


if __name__ == "__main__":
    build(verbose=True, extra_flags=sys.argv[1:])
