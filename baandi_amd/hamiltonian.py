"""Host-side mirror of the reference Gamma operator interface.

``GammaSystem`` is the set-up (refstart.rs:536-650), ``DistributedRealHamiltonian``
the operator (crates/dft-physics/src/distributed_real_gamma.rs:222-1264) with the
same method names and argument meaning: ``set_owned_potential(values, epoch)``,
``apply_block(x, y)`` (overwrites y), ``band_density_sum(x, weights)`` and
``nonlocal_band_energies(x)``. ORC blocks are torch tensors on the GPU with
shape (columns, rows): each tensor row is one contiguous ORC column, i.e. the
column-major n x m RealMatrix with leading dimension n.

All arithmetic runs in libbaandi_amd.so; torch supplies device memory only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L

# The reference's synthetic projector set: beta_i = r^(l+1) exp(-r^2/(i+1)),
# l = 0, 0, 1, D in Ry (crates/dft-engine/tests/common/mod.rs:163-172).
SYNTHETIC_BETAS = ((0, 1.0), (0, 2.0), (1, 3.0))
SYNTHETIC_DIJ_RY = ((2.0, 0.4, 0.0), (0.4, 4.0, 0.0), (0.0, 0.0, 6.0))


class GammaSystem:
    """One Gamma cell with one synthetic species, built natively."""

    def __init__(self, lattice_rows, frac_positions, z_valence, ecut_ha, rc=1.0, mesh_points=801,
                 betas=SYNTHETIC_BETAS, dij_ry=SYNTHETIC_DIJ_RY, grid_override=None, device=False):
        """``device``: run the set-up's atom loops on the GPU (bd_system_create_dev)."""
        lib = L.load()
        d = L.SystemDesc()
        for i in range(3):
            for j in range(3):
                d.lattice_rows[i][j] = float(lattice_rows[i][j])
        pos = np.ascontiguousarray(np.asarray(frac_positions, dtype=np.float64).reshape(-1))
        self._keep = [pos]
        d.natoms = len(pos) // 3
        d.frac_positions = pos.ctypes.data_as(C.POINTER(C.c_double))
        d.z_valence = float(z_valence)
        d.rc = float(rc)
        d.mesh_points = int(mesh_points)
        betas = list(betas or [])
        d.nproj = len(betas)
        if betas:
            pl = np.array([b[0] for b in betas], dtype=np.int32)
            pw = np.array([b[1] for b in betas], dtype=np.float64)
            dij = np.ascontiguousarray(np.asarray(dij_ry, dtype=np.float64).reshape(-1))
            self._keep += [pl, pw, dij]
            d.proj_l = pl.ctypes.data_as(C.POINTER(C.c_int32))
            d.proj_width = pw.ctypes.data_as(C.POINTER(C.c_double))
            d.dij_ry = dij.ctypes.data_as(C.POINTER(C.c_double))
        d.ecut_ha = float(ecut_ha)
        for i in range(3):
            d.grid_override[i] = 0 if grid_override is None else int(grid_override[i])
        h = C.c_void_p()
        create = lib.bd_system_create_dev if device else lib.bd_system_create
        L.check(create(C.byref(d), C.byref(h)))
        self.handle = h
        info = L.SystemInfo()
        L.check(lib.bd_system_get_info(h, C.byref(info)))
        self.n = info.n
        self.nrep = info.nrep
        self.dims = tuple(info.dims)
        self.nproj = info.nproj
        self.natoms = info.natoms
        self.ndensity = info.ndensity
        self.volume = info.volume
        self.electrons = info.electrons
        self.z_valence = float(z_valence)
        self.rc = float(rc)
        self.lattice_rows = [list(map(float, r)) for r in lattice_rows]

    def get(self, name):
        sizes = {
            "miller": (3 * self.n, np.int32), "g2": (self.n, np.float64),
            "kinetic_orc": (self.n, np.float64), "diagonal_static": (self.n, np.float64),
            "projectors_orc": (self.n * self.nproj, np.float64),
            "coupling": (self.nproj * self.nproj, np.float64),
            "plus_index": (self.nrep, np.int64), "minus_index": (self.nrep, np.int64),
            "plus_basis": (self.nrep, np.int64), "minus_basis": (self.nrep, np.int64),
            "density_miller": (3 * self.ndensity, np.int32), "density_g2": (self.ndensity, np.float64),
            "density_index": (self.ndensity, np.int64),
            "ionic_density_re": (self.ndensity, np.float64), "ionic_density_im": (self.ndensity, np.float64),
            "atoms_cart": (3 * self.natoms, np.float64), "ewald": (1, np.float64),
        }
        count, dtype = sizes[name]
        out = np.empty(count, dtype=dtype)
        L.check(L.load().bd_system_get(self.handle, name.encode(), out.ctypes.data_as(C.c_void_p), count))
        if name in ("miller", "density_miller", "atoms_cart"):
            out = out.reshape(-1, 3)
        if name == "projectors_orc":
            out = out.reshape(self.nproj, self.n).T
        if name == "coupling":
            out = out.reshape(self.nproj, self.nproj).T
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_system_destroy(h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self.handle = None


class DistributedRealHamiltonian:
    """The Gamma H·X operator on one GPU (DistributedRealOperator contract,
    crates/dft-solvers/src/distributed_davidson.rs:19-40)."""

    def __init__(self, system: GammaSystem, capacity, precision=L.BD_FP64):
        self.system = system
        self.capacity = int(capacity)
        h = C.c_void_p()
        L.check(L.load().bd_gamma_op_create(system.handle, self.capacity, int(precision), C.byref(h)))
        self.handle = h
        self.epoch = 0

    def global_len(self):
        return self.system.n

    def torch_stream(self):
        """The op's HIP stream as a torch stream object."""
        import torch
        if getattr(self, "_torch_stream", None) is None:
            self._torch_stream = torch.cuda.ExternalStream(self.stream())
        return self._torch_stream

    def enter(self):
        """Order the op's stream after the work torch has queued so far (the
        tensors handed to the op are complete before its kernels read them)."""
        import torch
        self.torch_stream().wait_stream(torch.cuda.current_stream())

    def leave(self):
        """Order torch's current stream after the op's queued work."""
        import torch
        torch.cuda.current_stream().wait_stream(self.torch_stream())

    def set_owned_potential(self, values, epoch=None):
        """V(r) on the full row-major grid (Ha): a CUDA float64 tensor or numpy array."""
        self.epoch = self.epoch + 1 if epoch is None else int(epoch)
        lib = L.load()
        if hasattr(values, "is_cuda") and values.is_cuda:
            v = values.contiguous()
            self.enter()
            L.check(lib.bd_gamma_op_set_potential_dev(self.handle, L.ptr(v), self.epoch))
            self.leave()
        else:
            v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
            L.check(lib.bd_gamma_op_set_potential(self.handle, L.ptr(v), self.epoch))

    def apply_block(self, x, y):
        """y = H x for CUDA float64 tensors of shape (m, n); y is overwritten.
        Ordered with torch's current stream on both sides."""
        m, n = x.shape
        assert n == self.global_len() and y.shape == x.shape and x.is_contiguous() and y.is_contiguous()
        self.enter()
        L.check(L.load().bd_gamma_op_apply_dev(self.handle, L.ptr(x), n, m, L.ptr(y), n))
        self.leave()

    def apply_block_host(self, x):
        """numpy (n, m) column-major convenience: returns H x."""
        x = np.asfortranarray(x, dtype=np.float64)
        y = np.empty_like(x, order="F")
        L.check(L.load().bd_gamma_op_apply(self.handle, L.ptr(x), x.shape[0], x.shape[1], L.ptr(y), x.shape[0]))
        return y

    def band_density_sum(self, x, weights, out):
        """out(grid) = sum_b w_b psi_b(r)^2 / volume for the rows of x."""
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
        self.enter()
        L.check(L.load().bd_gamma_op_band_density_dev(self.handle, L.ptr(x), x.shape[1], x.shape[0],
                                                       L.ptr(w), L.ptr(out)))
        self.leave()

    def keep_fields(self, on=True):
        """Keep the real-space pair fields of every later band density (the
        warm start's slabs; bd_gamma_op_keep_fields)."""
        L.check(L.load().bd_gamma_op_keep_fields(self.handle, 1 if on else 0))

    def apply_kept(self, x, delta_v, y):
        """y += dV x from the fields kept by the last band density of this
        same tensor x (apply_local_change, distributed_real_gamma.rs:1370-1438)."""
        m, n = x.shape
        assert n == self.global_len() and y.shape == x.shape and x.is_contiguous() and y.is_contiguous()
        dv = delta_v.reshape(-1).contiguous()
        self.enter()
        L.check(L.load().bd_gamma_op_apply_kept_dev(self.handle, L.ptr(x), n, m, L.ptr(dv), L.ptr(y), n))
        self.leave()

    def plane_split(self):
        """0 on the one-workgroup plane pass, else the split pass's z lines per batch."""
        v = np.zeros(1, dtype=np.int32)
        L.check(L.load().bd_gamma_op_plane_split(self.handle, L.ptr(v), 0))
        return int(v[0])

    def set_plane_split(self, zb):
        """0: one workgroup per plane (where it fits); zb > 0: the split pass."""
        v = np.array([zb], dtype=np.int32)
        L.check(L.load().bd_gamma_op_plane_split(self.handle, L.ptr(v), 1))

    def nonlocal_band_energies(self, x):
        out = np.empty(x.shape[0])
        self.enter()
        L.check(L.load().bd_gamma_op_nonlocal_energies_dev(self.handle, L.ptr(x), x.shape[1], x.shape[0],
                                                            L.ptr(out)))
        return out

    def stream(self):
        return L.load().bd_gamma_op_stream(self.handle)

    def autotune(self, columns, reps=3):
        """Time the valid launch shapes of the H·X kernels on `columns` columns
        of the op's own scratch and keep the fastest (bd_gamma_op_autotune):
        returns the 5 figures it reports. Every shape gives the same bits."""
        out = np.zeros(5)
        self.enter()
        L.check(L.load().bd_gamma_op_autotune(self.handle, int(columns), int(reps), L.ptr(out)))
        return tuple(out)

    def plane_shape(self):
        """(threads, y rows per z-FFT chunk) of the plane kernel."""
        out = np.zeros(2, dtype=np.int32)
        L.check(L.load().bd_gamma_op_plane_shape(self.handle, L.ptr(out), 0))
        return int(out[0]), int(out[1])

    def set_plane_shape(self, threads, chunk):
        """Force the plane kernel's shape (one the autotuner may pick)."""
        v = np.array([threads, chunk], dtype=np.int32)
        L.check(L.load().bd_gamma_op_plane_shape(self.handle, L.ptr(v), 1))

    def set_profiling(self, on=True):
        L.check(L.load().bd_gamma_op_set_profiling(self.handle, 1 if on else 0))

    def plan_info(self):
        out = np.zeros(8, dtype=np.int64)
        L.check(L.load().bd_gamma_op_plan_info(self.handle, L.ptr(out)))
        keys = ("n0", "n1", "n2", "sticks", "sticks_padded", "z_planes", "tile", "tiles")
        return {k: int(v) for k, v in zip(keys, out)}

    def stage_totals(self):
        """Cumulative event timing since set_profiling(True): dict of ms per
        stage and the launch/pair/column counts."""
        out = np.zeros(8)
        L.check(L.load().bd_gamma_op_stage_times(self.handle, L.ptr(out)))
        names = ("stick_inverse", "plane", "stick_forward", "projections", "back_product")
        d = {k: float(v) for k, v in zip(names, out[:5])}
        d.update(applies=int(out[5]), pairs=int(out[6]), columns=int(out[7]))
        return d

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                L.load().bd_gamma_op_destroy(h)
            except (TypeError, AttributeError):  # interpreter shutdown
                pass
            self.handle = None
