"""Build the C-ABI shared library (hierarchicaldd_amd/_build/libhdd.so) with hipcc for gfx950.

Every source is compiled by hipcc (host .cpp files too, so one toolchain links
the HIP runtime). Floating-point contraction is off so the kernels keep the
reference's operation order, and packed-fp32 instruction selection is
disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libhdd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I", os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    # host-only sources: AVX2/FMA instructions for the setup's dense kernels (no FP contraction, so results are unchanged)
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-march=x86-64-v3"]
    cmd = [HIPCC] + COMMON + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


KRYLOV_CHECK_SRC = os.path.join(HERE, "..", "tests", "krylov_adapter_check.cpp")
KRYLOV_CHECK_LIB = os.path.join(OUT, "libhdd_krylov_check.so")


def build_krylov_check() -> str:
    """Compiles the test driver of include/hdd_krylov.hpp (a host-only C++ file that
    links libhdd.so) into _build/libhdd_krylov_check.so."""
    tmp = KRYLOV_CHECK_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(HERE, "..", "include"), KRYLOV_CHECK_SRC,
           "-o", tmp, "-L", OUT, "-lhdd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("krylov adapter check failed to build:\n" + r.stderr[-4000:])
    os.replace(tmp, KRYLOV_CHECK_LIB)
    return KRYLOV_CHECK_LIB


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-fopenmp", "--offload-arch=gfx950", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    build_krylov_check()
    if verbose:
        print("built", LIB, "and", KRYLOV_CHECK_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
