"""Build libmfp.so (HIP, gfx950) in-tree.

``python -m matrix_format_performance_amd.build`` compiles every source under
``csrc/`` with hipcc (in parallel) and links ``libmfp.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmfp.so")
APP = os.path.join(HERE, "mfp_bench")
RS_SLOT = os.path.join(HERE, "librocsparse_slot.so")

SOURCES = [
    "runtime.hip",
    "convert.hip",
    "spmv.hip",
    "plan.hip",
    "reference_api.hip",
    "precision.hip",
    "batch.hip",
    "features.hip",
    "mlp.hip",
    "autotune.cpp",
    "generate.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # explicit fma only: keeps the kernels' summation orders reproducible
    "-ffp-contract=off",
    # hardware floating-point atomics (global_atomic_add_f32/f64, ds_add)
    "-munsafe-fp-atomics",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile and link libmfp.so; returns its path."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; libmfp needs ROCm to build")
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_app()
    if verbose:
        print(f"built {LIB} and {APP}")
    return LIB


def build_app() -> str:
    """The C++ benchmark driver (app/mfp_bench.cpp) over the C ABI and the
    reference-compatible header include/mfp_compat.hpp; host code only, so
    g++ builds it (x86-64-v3 so it runs on any AVX2 host)."""
    cxx = os.environ.get("CXX", "g++")
    tmp = APP + ".tmp"
    cmd = [cxx, "-O3", "-std=c++17", "-march=x86-64-v3", "-Wall", "-Wextra", "-Wno-unused-parameter",
           os.path.join(HERE, "app", "mfp_bench.cpp"), os.path.join(HERE, "app", "rocsparse_slot.cpp"), "-o", tmp,
           "-L" + HERE, "-lmfp", "-Wl,-rpath,$ORIGIN", "-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mfp_bench build failed:\n{r.stderr}")
    os.replace(tmp, APP)
    build_rocsparse_slot()
    return APP


def build_rocsparse_slot() -> str:
    """The benchmark-only rocSPARSE comparator (app/rocsparse_slot.cpp) as a
    shared library for bench.py; it opens librocsparse.so at run time."""
    cxx = os.environ.get("CXX", "g++")
    tmp = RS_SLOT + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
           os.path.join(HERE, "app", "rocsparse_slot.cpp"), "-o", tmp, "-L" + HERE, "-lmfp",
           "-Wl,-rpath,$ORIGIN", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rocSPARSE slot build failed:\n{r.stderr}")
    os.replace(tmp, RS_SLOT)
    return RS_SLOT


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
