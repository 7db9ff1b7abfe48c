"""Builds ddla_amd/libddla_hip.so from ddla_amd/csrc/*.hip with hipcc for gfx950.

Usage: python -m ddla_amd.build [--jobs N] [--incremental]
Objects go to ddla_amd/_build/; the shared library is written in-tree next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libddla_hip.so")
ARCH = os.environ.get("DDLA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (only scalar / vector forms the compiler proves safe)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Per-source flags.  The LU panel kernels keep up to 64 rows per lane in register arrays indexed
# by unrolled constants; AMDGPU's promote-alloca pass would merge each array into one wide
# vector value, and every uniform branch then copies the whole vector.  Left to SROA, the
# elements stay independent registers.
FILE_FLAGS = {
    "lu_kernels.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],
    # same for the Cholesky diagonal-block kernel's per-lane row of the 32-column block
    "chol_kernels.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],
    # the batched kernels unroll every column step of a panel (register-indexed rows), which
    # exceeds the default size limit for fully unrolled loops
    "batched.hip": ["-mllvm", "-pragma-unroll-threshold=200000"],
}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src, os.path.join(os.path.dirname(HERE), "include", "ddla_native.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *CFLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> str:
    """Compiles every source (incremental=True: only sources newer than their objects or
    than the headers) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4) // 2))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources changed since their objects were built")
    a = ap.parse_args()
    build(a.jobs, incremental=a.incremental)
