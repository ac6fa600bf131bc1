"""Builds the gfx950 shared library ``disco_amd/_lib/libdisco_gpu.so``.

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instruction selection disabled (``-packed-fp32-ops``).  The host set-up code
is compiled as HIP too so that the physics leaves it shares with the kernels
are the same ``__host__ __device__`` functions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libdisco_gpu.so"

SOURCES = ["host_setup.cpp", "checkpoint.cpp", "hydro_step.hip", "capi.hip", "c2p.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


def hip_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-function",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", CSRC,
        "-I", os.path.join(os.path.dirname(HERE), "include"),
    ]


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False) -> str:
    """Compile all sources (every call recompiles) and link the library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    # the translation units compile concurrently (hydro_step.hip, with every
    # physics module's kernels, takes most of the time)
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-x", "hip", *hip_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = lib_path() + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib_path())
    shutil.rmtree(objdir, ignore_errors=True)
    return lib_path()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
