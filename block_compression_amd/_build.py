"""Builds the HIP library block_compression_amd/_lib/libbcn_mi355x.so.

hipcc compiles every csrc/*.hip for gfx950 into one shared library with a C
ABI (include/bcn_mi355x.h).  The flags carry the numeric contract of the
kernels: no FMA contraction, correctly rounded f32 division and square root,
IEEE denormals, and no packed-FP32 instructions.
"""
from __future__ import annotations

import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libbcn_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
]


# Per-source code-generation flags.  The BC7 encoders issue ~130 K VALU
# instructions per block and are bound by VALU issue at three waves per SIMD;
# LLVM's post-register-allocation scheduler reorders their straight-line code
# for latency at the cost of issue slots, and the 4096x4096 opaque_basic
# launch runs ~1% faster without it (same instructions, same results).
# Hazard wait states are inserted by a separate pass either way.
FILE_FLAGS = {
    "bc7_encode.hip": ["-mllvm", "-disable-post-ra"],
    "batch_encode.hip": ["-mllvm", "-disable-post-ra"],
}


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def _compile(src: str, obj: str, verbose: bool) -> None:
    cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], *FILE_FLAGS.get(os.path.basename(src), []),
           "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles each source to an object in parallel (one hipcc per file),
    then links the shared library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(obj_dir, os.path.basename(s) + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_compile, s, o, verbose) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB_PATH)
