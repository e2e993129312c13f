"""Build libbaandi_amd.so in-tree for gfx950 with hipcc.

    python -m baandi_amd.build        # or build() from __graft_entry__

Every source under baandi_amd/csrc is compiled into one shared library next
to this file. HIP code is built with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbaandi_amd.so")
ARCH = "gfx950"

SOURCES = ["errors.cpp", "setup.cpp", "gemm.hip", "gemm_bf16x3.hip", "op.hip", "kbatch.hip", "eigen.hip", "subspace.hip",
           "mlp.hip", "davidson.hip", "kdavidson.hip", "field.hip", "setup_dev.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off"]


def _hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP extension cannot be built")


def build(verbose=False, jobs=8):
    hipcc = _hipcc()
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        path = os.path.join(CSRC, src)
        obj = os.path.join(obj_dir, src + ".o")
        if src.endswith(".hip"):
            cmd = [hipcc, *HIP_FLAGS, "-c", path, "-o", obj]
        else:
            cmd = ["g++", *HOST_FLAGS, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose and (r.stderr.strip()):
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
