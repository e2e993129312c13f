"""CPU FP64 H·X with OpenMP (the CPU baseline's timed restatement).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): bench.py's
``cpu_baseline`` leg times it, tests/ check it against ``oracle.hamiltonian``.

The local part (pair packing, pruned stick/plane FFTs, V(r), unpack, kinetic)
is ``cpu_hx.cpp`` built with g++ -O3 -fopenmp into ``oracle/_build``; the
nonlocal products P D P^T X are NumPy BLAS GEMMs, as the reference does them
with faer or MKL dgemm (crates/dft-physics/src/distributed_real_gamma.rs:957-1251).

Parity: equal to ``oracle.hamiltonian.apply_orc`` to 1e-12 relative on the
KAT-1, Si2 and KAT-2 cells (tests/test_oracle_cpu_hx.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libbd_cpu_hx.so")
SRC = os.path.join(HERE, "cpu_hx.cpp")


def build():
    """Compile cpu_hx.cpp (every call)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fcx-limited-range", "-fPIC", "-shared", "-fopenmp", SRC,
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle CPU H·X build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P, I64 = C.c_void_p, C.c_int64
        lib.bd_cpu_hx_create.argtypes = [C.c_int, C.c_int, C.c_int, I64, I64, P, P, P]
        lib.bd_cpu_hx_create.restype = P
        lib.bd_cpu_hx_destroy.argtypes = [P]
        lib.bd_cpu_hx_set_potential.argtypes = [P, P]
        lib.bd_cpu_hx_set_threads.argtypes = [P, C.c_int]
        lib.bd_cpu_hx_plan.argtypes = [P, P]
        lib.bd_cpu_hx_apply_local.argtypes = [P, P, I64, C.c_int, P, I64]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class CpuHamiltonian:
    """H·X of one Gamma cell on the host cores. Built from the native set-up
    arrays of ``baandi_amd.hamiltonian.GammaSystem`` (plus/minus FFT indices,
    T_G per ORC row, ORC projectors, per-atom D blocks)."""

    def __init__(self, system, threads=None):
        lib = load()
        self.n = system.n
        self.dims = tuple(system.dims)
        self._plus = np.ascontiguousarray(system.get("plus_index"), dtype=np.int64)
        self._minus = np.ascontiguousarray(system.get("minus_index"), dtype=np.int64)
        self._kin = np.ascontiguousarray(system.get("kinetic_orc"))
        self.h = lib.bd_cpu_hx_create(self.dims[0], self.dims[1], self.dims[2], self.n, len(self._plus),
                                      _p(self._plus), _p(self._minus), _p(self._kin))
        self.proj = np.asfortranarray(system.get("projectors_orc"))
        d = system.get("coupling")
        nproj = d.shape[0]
        nc = nproj // system.natoms if system.natoms else 0  # channels per atom
        self.blocks = [(s, s + nc, np.ascontiguousarray(d[s:s + nc, s:s + nc])) for s in range(0, nproj, nc or 1)]
        self.threads = threads or os.cpu_count() or 1
        self.set_threads(self.threads)

    def set_threads(self, threads):
        self.threads = int(threads)
        load().bd_cpu_hx_set_threads(self.h, self.threads)

    def plan(self):
        out = np.zeros(2, dtype=np.int64)
        load().bd_cpu_hx_plan(self.h, _p(out))
        return {"sticks": int(out[0]), "z_planes": int(out[1])}

    def set_potential(self, v):
        self._v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        load().bd_cpu_hx_set_potential(self.h, _p(self._v))

    def apply(self, x):
        """Y = H X for an (n, m) ORC block (Fortran or C order accepted)."""
        x = np.asfortranarray(x, dtype=np.float64)
        n, m = x.shape
        y = np.empty((n, m), order="F")
        load().bd_cpu_hx_apply_local(self.h, _p(x), n, m, _p(y), n)
        if self.proj.shape[1]:
            q = self.proj.T @ x
            w = np.empty_like(q)
            for s, e, d in self.blocks:
                w[s:e] = d @ q[s:e]
            y += self.proj @ w
        return y

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                load().bd_cpu_hx_destroy(h)
            except (TypeError, AttributeError):
                pass
            self.h = None
