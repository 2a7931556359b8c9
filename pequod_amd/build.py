"""Build libpequod_hip.so (HIP kernels + C ABI) for gfx950, in-tree.

Every source is compiled with hipcc directly (no torch JIT, no cmake): the
resulting shared library sits next to this file so that it travels with the
source tree.  Packed-FP32 instructions are disabled for device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpequod_hip.so")
ARCH = os.environ.get("PQ_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "pq_runtime.cpp",
    "pq_inversion.hip",
    "pq_cabaret.hip",
    "pq_ensemble.hip",
    "pq_diag.hip",
    "pq_io.cpp",
    "pq_serve.cpp",
]

# Transform lengths N = g - 1 with a kernel unit each (pq_inv_kernels.hip
# compiled with -DPQ_N=N); must match PQ_INV_SIZES in csrc/pq_inv.h.
INV_SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024,
             1536, 2048, 3072, 4096]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(job) -> str:
    src, n = job if isinstance(job, tuple) else (job, None)
    path = os.path.join(CSRC, src)
    stem = os.path.splitext(src)[0] + ("" if n is None else f"_n{n}")
    obj = os.path.join(BUILD, stem + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, "-c", path, "-o", obj]
    if n is not None:
        cmd.insert(1, f"-DPQ_N={n}")
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libpequod_hip.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    # largest units first so that the pool stays busy
    jobs += [("pq_inv_kernels.hip", n) for n in sorted(INV_SIZES, reverse=True)]
    workers = max(1, min(len(jobs), (os.cpu_count() or 4), 8))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libpequod_hip.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
