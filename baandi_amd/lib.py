"""ctypes binding of libbaandi_amd.so (the C ABI in include/baandi_amd.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbaandi_amd.so")

BD_OK, BD_ERR_LINALG, BD_ERR_INVALID_PARAMETER, BD_ERR_APPLICATION, BD_NOT_CONVERGED, BD_ERR_DEVICE = range(6)
BD_FP64, BD_MIXED, BD_FP32_HX = 0, 1, 2
BD_PRECOND_QE, BD_PRECOND_TPA, BD_PRECOND_MLP = 0, 1, 2

# Every function include/baandi_amd.h declares.
EXPORTS = [
    "bd_last_error", "bd_system_create", "bd_system_create_dev", "bd_system_destroy", "bd_system_get_info", "bd_system_get",
    "bd_gamma_op_create", "bd_gamma_op_destroy", "bd_gamma_op_set_potential",
    "bd_gamma_op_set_potential_dev", "bd_gamma_op_apply", "bd_gamma_op_apply_dev",
    "bd_gamma_op_band_density_dev", "bd_gamma_op_nonlocal_energies_dev", "bd_gamma_op_stream",
    "bd_gamma_op_set_profiling", "bd_gamma_op_stage_times", "bd_gamma_op_plan_info", "bd_davidson_solve_dev",
    "bd_davidson_solve_many_dev", "bd_k_band_density_dev", "bd_k_set_potential_dev", "bd_k_davidson_create", "bd_k_davidson_destroy", "bd_k_davidson_solve_dev", "bd_kblocks_create", "bd_kblocks_destroy", "bd_kblocks_set_diagonals_dev", "bd_kblocks_start_dev", "bd_kblocks_step", "bd_kblocks_vectors_dev", "bd_block_create", "bd_block_destroy", "bd_block_set_diagonal_dev",
    "bd_block_start_dev", "bd_block_step", "bd_block_step_async", "bd_block_finish", "bd_block_vectors_dev",
    "bd_k_op_create", "bd_k_op_info", "bd_k_op_get", "bd_k_batch_create", "bd_k_batch_destroy",
    "bd_k_batch_stream", "bd_k_batch_apply_dev", "bd_gamma_op_autotune", "bd_gamma_op_fused_shape",
    "bd_gamma_op_plane_shape", "bd_gamma_op_plane_split", "bd_gamma_op_keep_fields", "bd_gamma_op_apply_kept_dev",
    "bd_sym_eigen_multi_from", "bd_herm_eigen_dev", "bd_sym_eigen_coop", "bd_field_create", "bd_field_destroy", "bd_field_stream", "bd_field_to_grid_dev",
    "bd_field_from_grid_dev", "bd_field_potential_dev", "bd_field_residual_dev", "bd_field_gram_dev",
    "bd_field_mix_dev", "bd_field_energies_dev",
    "bd_gemm_workspace", "bd_gemm_dev", "bd_gemm_time_dev", "bd_mfma_probe_dev", "bd_mfma_f64_probe_dev",
    "bd_sym_eigen_workspace", "bd_sym_eigen_dev", "bd_inverse_cholesky_dev", "bd_inverse_cholesky_static_dev",
    "bd_mlp_param_count", "bd_mlp_blob_floats", "bd_mlp_quantize_dev", "bd_mlp_forward_dev",
    "bd_mlp_train_workspace", "bd_mlp_loss_grad_dev", "bd_mlp_adam_dev", "bd_mlp_train_dev",
]


class SystemDesc(C.Structure):
    _fields_ = [
        ("lattice_rows", (C.c_double * 3) * 3),
        ("natoms", C.c_int32),
        ("frac_positions", C.POINTER(C.c_double)),
        ("z_valence", C.c_double),
        ("rc", C.c_double),
        ("mesh_points", C.c_int32),
        ("nproj", C.c_int32),
        ("proj_l", C.POINTER(C.c_int32)),
        ("proj_width", C.POINTER(C.c_double)),
        ("dij_ry", C.POINTER(C.c_double)),
        ("ecut_ha", C.c_double),
        ("grid_override", C.c_int32 * 3),
    ]


class SystemInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nrep", C.c_int64), ("dims", C.c_int32 * 3), ("nproj", C.c_int64),
        ("natoms", C.c_int64), ("ndensity", C.c_int64), ("volume", C.c_double), ("electrons", C.c_double),
    ]


class DavidsonOptions(C.Structure):
    _fields_ = [
        ("bands", C.c_int32), ("threshold", C.c_double), ("occupied", C.c_int32),
        ("first_step", C.c_int32), ("capacity", C.c_int32), ("max_iterations", C.c_int32),
        ("return_hx", C.c_int32), ("preconditioner", C.c_int32),
        ("mlp_blob", C.c_void_p), ("mlp_vbar", C.c_double),
        ("profile", C.c_int32), ("ritz_changes_ha", C.c_void_p), ("ritz_values_ha", C.c_void_p),
        ("complex_columns", C.c_int32),
    ]


class DavidsonRecord(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("h_calls", C.c_int32), ("h_columns", C.c_int32),
        ("repairs", C.c_int32), ("restarts", C.c_int32), ("all_converged", C.c_int32),
        ("seconds_total", C.c_double), ("seconds_apply", C.c_double), ("seconds_dense", C.c_double),
        ("mlp_fallback", C.c_int32), ("eigen_retries", C.c_int32),
        ("ideal_resets", C.c_int32), ("periodic_refreshes", C.c_int32), ("guard_fallbacks", C.c_int32),
    ]


class BaandiError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m baandi_amd.build` (no CPU fallback)")
    # torch ships its own libamdhip64.so.7. Loading it first makes the library
    # bind to that same runtime (one HIP runtime per process); loaded the other
    # way round, torch finds no device behind the system runtime.
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, D, U64, F = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64, C.c_float
    dp = C.POINTER(C.c_double)
    lib.bd_last_error.restype = C.c_char_p
    sig = {
        "bd_system_create": [C.POINTER(SystemDesc), C.POINTER(P)],
        "bd_system_create_dev": [C.POINTER(SystemDesc), C.POINTER(P)],
        "bd_system_get_info": [P, C.POINTER(SystemInfo)],
        "bd_system_get": [P, C.c_char_p, P, I64],
        "bd_gamma_op_create": [P, I32, I32, C.POINTER(P)],
        "bd_k_op_create": [P, P, I32, I32, C.POINTER(P)],
        "bd_k_op_info": [P, P, P, P],
        "bd_k_op_get": [P, C.c_char_p, P, I64],
        "bd_k_batch_create": [P, I32, C.POINTER(P)],
        "bd_k_batch_apply_dev": [P, P, P, P, P, P],
        "bd_gamma_op_autotune": [P, I32, I32, P],
        "bd_gamma_op_fused_shape": [P, P, I32],
        "bd_gamma_op_plane_shape": [P, P, I32],
        "bd_gamma_op_plane_split": [P, P, I32],
        "bd_gamma_op_keep_fields": [P, I32],
        "bd_gamma_op_apply_kept_dev": [P, P, I64, I32, P, P, I64],
        "bd_sym_eigen_multi_from": [P, I32],
        "bd_sym_eigen_coop": [P, I32],
        "bd_field_create": [P, C.POINTER(P)],
        "bd_field_to_grid_dev": [P, P, P],
        "bd_field_from_grid_dev": [P, P, P],
        "bd_field_potential_dev": [P, P, P, P, P],
        "bd_field_residual_dev": [P, P, P, P, P],
        "bd_field_gram_dev": [P, I32, P, P, P, P],
        "bd_field_mix_dev": [P, I32, P, P, P, P, P, D, I32, P],
        "bd_field_energies_dev": [P, P, P, P],
        "bd_gamma_op_set_potential": [P, P, U64],
        "bd_gamma_op_set_potential_dev": [P, P, U64],
        "bd_gamma_op_apply": [P, P, I64, I32, P, I64],
        "bd_gamma_op_apply_dev": [P, P, I64, I32, P, I64],
        "bd_gamma_op_band_density_dev": [P, P, I64, I32, P, P],
        "bd_gamma_op_nonlocal_energies_dev": [P, P, I64, I32, P],
        "bd_gamma_op_set_profiling": [P, I32],
        "bd_gamma_op_stage_times": [P, P],
        "bd_gamma_op_plan_info": [P, P],
        "bd_davidson_solve_dev": [P, P, I64, P, P, C.POINTER(DavidsonOptions), P, P, P, I64, P,
                                  C.POINTER(DavidsonRecord)],
        "bd_davidson_solve_many_dev": [I32, P, P, P, P, C.POINTER(DavidsonOptions), P, P, P, P, P,
                                       C.POINTER(DavidsonRecord), P, I32],
        "bd_k_band_density_dev": [P, I32, P, P, P, P, P, P],
        "bd_k_set_potential_dev": [P, I32, P, U64, P],
        "bd_k_davidson_create": [P, I32, I32, I32, C.POINTER(P)],
        "bd_k_davidson_solve_dev": [P, P, P, P, P, C.POINTER(DavidsonOptions), P, P, P, P, P,
                                    C.POINTER(DavidsonRecord), P],
        "bd_kblocks_create": [P, I32, I32, I32, C.POINTER(P)],
        "bd_kblocks_set_diagonals_dev": [P, P],
        "bd_kblocks_start_dev": [P, P, P, P],
        "bd_kblocks_step": [P, I32, I32, P],
        "bd_kblocks_vectors_dev": [P, P, P, P],
        "bd_block_create": [P, I32, I32, C.POINTER(P)],
        "bd_block_set_diagonal_dev": [P, P],
        "bd_block_start_dev": [P, P, I64, P],
        "bd_block_step": [P, I32, I32, P],
        "bd_block_step_async": [P, I32, I32],
        "bd_block_finish": [P, P],
        "bd_block_vectors_dev": [P, P, P, I64],
        "bd_gemm_dev": [I32, I32, I64, I64, I64, D, P, I64, P, I64, D, P, I64, P],
        "bd_gemm_time_dev": [I32, I32, I64, I64, I64, P, I64, P, I64, P, I64, P, I32, P],
        "bd_mfma_probe_dev": [P, P, P],
        "bd_mfma_f64_probe_dev": [P, P, P],
        "bd_sym_eigen_dev": [I32, I32, P, I64, P, P, I64, P],
        "bd_herm_eigen_dev": [I32, I32, P, I64, P, P, I64, P],
        "bd_inverse_cholesky_dev": [P, P, I32, I32, I32, P, P],
        "bd_inverse_cholesky_static_dev": [P, P, I32, I32, I32, P, P],
        "bd_mlp_quantize_dev": [P, P],
        "bd_mlp_forward_dev": [I64, P, F, F, F, P, P],
        "bd_mlp_loss_grad_dev": [I64, I32, P, P, P, P, P, I32, P, P, P, P],
        "bd_mlp_adam_dev": [P, P, P, P, I32, F],
        "bd_mlp_train_dev": [I64, I32, P, P, P, P, P, P, P, I32, I32, F, U64, P, P, P, P, P],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.bd_system_destroy.argtypes = [P]
    lib.bd_system_destroy.restype = None
    lib.bd_gamma_op_destroy.argtypes = [P]
    lib.bd_gamma_op_destroy.restype = None
    lib.bd_gamma_op_stream.argtypes = [P]
    lib.bd_gamma_op_stream.restype = P
    lib.bd_block_destroy.argtypes = [P]
    lib.bd_block_destroy.restype = None
    lib.bd_kblocks_destroy.argtypes = [P]
    lib.bd_kblocks_destroy.restype = None
    lib.bd_k_davidson_destroy.argtypes = [P]
    lib.bd_k_davidson_destroy.restype = None
    lib.bd_field_destroy.argtypes = [P]
    lib.bd_field_destroy.restype = None
    lib.bd_field_stream.argtypes = [P]
    lib.bd_field_stream.restype = P
    lib.bd_k_batch_destroy.argtypes = [P]
    lib.bd_k_batch_destroy.restype = None
    lib.bd_k_batch_stream.argtypes = [P]
    lib.bd_k_batch_stream.restype = P
    lib.bd_gemm_workspace.argtypes = [I32, I32, I64, I64, I64]
    lib.bd_gemm_workspace.restype = I64
    lib.bd_sym_eigen_workspace.argtypes = [I32, I32]
    lib.bd_sym_eigen_workspace.restype = I64
    lib.bd_mlp_param_count.argtypes = []
    lib.bd_mlp_param_count.restype = I32
    lib.bd_mlp_blob_floats.argtypes = []
    lib.bd_mlp_blob_floats.restype = I32
    lib.bd_mlp_train_workspace.argtypes = [I64, I32]
    lib.bd_mlp_train_workspace.restype = I64
    _lib = lib
    return lib


def check(status, allow=()):
    if status != BD_OK and status not in allow:
        msg = load().bd_last_error().decode()
        raise BaandiError(status, msg)
    return status


def ptr(a):
    """Address of a numpy array or torch tensor as void*."""
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)
