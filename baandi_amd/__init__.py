"""baandi_amd: MI355X-native Davidson H·ψ path of baandi (gfx950 HIP kernels
behind a C ABI, include/baandi_amd.h)."""
from .lib import load, BaandiError, BD_FP64, BD_MIXED, BD_FP32_HX, BD_PRECOND_QE, BD_PRECOND_TPA  # noqa: F401
