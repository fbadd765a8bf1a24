"""Compile the gfx950 library in-tree: atrbvh_amd/libatrbvh.so.

Every HIP translation unit is built for gfx950 only, with FMA contraction off
(the SAH / area arithmetic is part of the bit-exact contract), correctly
rounded fp32 division and square root, and packed-fp32 instruction selection
disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatrbvh.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("ATRBVH_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]

SOURCES = ["lbvh.hip", "treelet.hip", "collapse.hip", "trace.hip", "batch.hip", "adjoint.hip",
           "noalloc.hip", "compact.hip", "raysort.hip", "api.cpp", "obj.cpp",
           "scenes.cpp", "schedule.cpp"]

# Per-file flags.  treelet.hip only ever handles finite boxes, areas and SAH
# values (scenes reject non-finite vertices), so fminf/fmaxf there need no
# NaN quieting of their operands: without it the agglomeration loop issues
# ~12% fewer VALU instructions, with identical results.  The traversal keeps
# IEEE NaN handling (its slab tests can form 0 * inf).
EXTRA = {"treelet.hip": ["-fno-honor-nans"]}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, *EXTRA.get(src, [])]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    noise = [ln for ln in r.stderr.splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "unused-command-line" not in ln]
    if noise:
        sys.stderr.write(f"[{src}]\n" + "\n".join(noise) + "\n")
    return obj


def build(jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_plain_loop()
    return LIB


PLAIN_LOOP = os.path.join(HERE, "plain_loop")


def build_plain_loop() -> str:
    """The C++ per-mesh loop the benchmark times beside the batched build
    (host code only, against the library's C ABI)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", os.path.join(HERE, "plain_loop.cpp"),
           "-L" + HERE, "-latrbvh", "-Wl,-rpath,$ORIGIN", "-o", PLAIN_LOOP]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"plain_loop build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return PLAIN_LOOP


if __name__ == "__main__":
    print(build())
