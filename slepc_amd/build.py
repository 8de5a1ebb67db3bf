"""Builds the two shared libraries of the package, in-tree:

* ``slepc_amd/libks_hip.so``   -- HIP kernels + C ABI (include/ks_hip.h), gfx950
* ``slepc_amd/libslepc_amd.so`` -- the C host framework (include/slepc_amd.h),
  linked against libks_hip.so and MKL's LAPACK/BLAS.

Run ``python -m slepc_amd.build``.  hipcc cross-compiles gfx950 code without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
MKLDIR = os.environ.get("MKLDIR", "/opt/conda/lib")
ARCH = "gfx950"

HIP_SOURCES = ["ks_runtime.hip", "ks_spmv.hip", "ks_bv.hip", "ks_group.hip", "ks_transpose.hip", "ks_ksp.hip", "ks_tsqr.hip", "ks_lu.hip"]
HOST_SOURCES = ["sys.c", "vec.c", "mat.c", "bv.c", "bvhipks.c", "ds.c", "st.c", "eps.c", "krylovschur.c", "serve.c", "matload.c"]

KS_LIB = os.path.join(HERE, "libks_hip.so")
HOST_LIB = os.path.join(HERE, "libslepc_amd.so")


def _run(cmd, quiet_filter=None):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = p.stdout
    if quiet_filter:
        out = "\n".join(l for l in out.splitlines() if quiet_filter not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build_kernels():
    src = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    tmp = KS_LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           # packed FP32 instructions stay off on these parts
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-o", tmp] + src
    _run(cmd, quiet_filter="packed-fp32-ops")
    os.replace(tmp, KS_LIB)


def build_host():
    src = [os.path.join(HERE, "host", s) for s in HOST_SOURCES]
    tmp = HOST_LIB + ".tmp"
    cmd = [CC, "-O2", "-g", "-fPIC", "-shared", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-Wall",
           "-Wno-unused-function", "-o", tmp] + src + [
           f"-L{HERE}", "-lks_hip", "-Wl,-rpath,$ORIGIN",
           f"-L{MKLDIR}", f"-Wl,-rpath,{MKLDIR}", "-lmkl_rt", "-lm"]
    _run(cmd)
    os.replace(tmp, HOST_LIB)


GEN_LIB = os.path.join(HERE, "libks_gen.so")


def build_gen():
    """Workload generators (R-MAT graph Laplacian) with OpenMP."""
    tmp = GEN_LIB + ".tmp"
    _run([CC, "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall", "-o", tmp,
          os.path.join(HERE, "gen", "rmat.c"), "-lm"])
    os.replace(tmp, GEN_LIB)


def build_oracle():
    if shutil.which("make"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(oracle: bool = True):
    build_kernels()
    build_host()
    build_gen()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", KS_LIB, HOST_LIB)
