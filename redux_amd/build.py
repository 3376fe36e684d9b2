"""Builds redux_amd/_lib/libmomfbd_hip.so in-tree for gfx950.

Device code (csrc/kernels.hip, csrc/mosaic.hip, csrc/basis.hip) is compiled with hipcc; the host solver and setup
(C++20) with g++; hipcc links the shared library.  Object files go to
redux_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libmomfbd_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = ["-O3", "-std=c++20", "-fPIC", "--offload-arch=gfx950", "-Wall", "-Wno-unused-function",
             # packed FP32 instructions are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # matrix-core accumulators in ordinary VGPRs (gfx950's register file is unified):
             # without it the mode GEMM copies its accumulators between VGPRs and AGPRs on
             # every k-chunk (FP64 k_phase_mfma 113 -> 105 ms per 256-patch step, same results)
             "-mllvm", "-amdgpu-mfma-vgpr-form"]
CXX_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{ROCM}/include",
             "-D__HIP_PLATFORM_AMD__",
             # complex products as (ac - bd, ad + bc) without the NaN/Inf recovery call
             "-fcx-limited-range"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[0])} (exit {r.returncode})")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    jobs = []
    for name in ("kernels.hip", "mosaic.hip", "basis.hip"):
        obj = os.path.join(BUILD, name + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, name), "-o", obj])
        objs.append(obj)
    for name in ("solver.cpp", "setup.cpp", "writers.cpp"):
        obj = os.path.join(BUILD, name + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, name), "-o", obj])
        objs.append(obj)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for j in jobs]
    for j, p in zip(jobs, procs):
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(j) + "\n" + out + err)
            raise RuntimeError(f"compile failed: {j[-3]}")
        if verbose and (out or err):
            sys.stderr.write(out + err)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
