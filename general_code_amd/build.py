"""Build the HIP library libmag_gpu.so (gfx950) in-tree.

    python -m general_code_amd.build          # or build() from Python

Each .hip source is compiled to an object in parallel, then linked into
general_code_amd/libmag_gpu.so, which the Python front end and the C++
host code load.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmag_gpu.so")
SOURCES = ["mag_sweep.hip", "mag_sweep64.hip", "mag_kernels.hip", "mag_generate.hip", "mag_spin.hip",
           "mag_capi.hip", "mag_spin_capi.hip", "mag_batch.hip", "mag_server.hip",
           "mag_tune.hip", "mag_monopole.hip", "mag_noalloc.hip"]
ARCH = os.environ.get("MAG_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-command-line-argument",
]


# Sources built without floating-point contraction: the FP64 link-form
# half-sweeps and the layout / rebuild kernels, whose FP64 results then
# follow the source's IEEE operations exactly (as the oracle's C code, also
# built with -ffp-contract=off, does), bit for bit.
NO_CONTRACT = {"mag_sweep64.hip", "mag_kernels.hip"}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    extra = ["-ffp-contract=off"] if src in NO_CONTRACT else []
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cli(verbose)
    return LIB


CLI = os.path.join(PKG, "mag_fixation")


def build_cli(verbose: bool = False) -> str:
    """The mag_fixation command-line app (host C++ over the C ABI)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(PKG, "cpp"), os.path.join(PKG, "cpp", "mag_fixation.cpp"),
           "-L", PKG, "-lmag_gpu", "-Wl,-rpath,$ORIGIN", "-o", CLI]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mag_fixation build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {CLI}")
    return CLI


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
