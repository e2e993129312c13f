// Wave-wide reductions with DPP row operations (GCN/CDNA cross-lane moves:
// row_shr within rows of 16 lanes, then row_bcast:15 and row_bcast:31 across
// rows), which stay in the VALU instead of going through the LDS crossbar
// like __shfl_xor. The sum is formed in lane 63 in a fixed order and read
// back to every lane.
#pragma once
#include <hip/hip_runtime.h>

namespace bd {

namespace dpp {
// DPP control words (GFX9): row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double move(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, BANK_MASK, true);
  const int rhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, BANK_MASK, true);
  return __hiloint2double(rhi, rlo);
}
}  // namespace dpp

// Sum of v over the 64 lanes of the wave, returned to every lane.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp::move<0x111, 0xF, 0xF>(v);  // row_shr:1 (inclusive scan within each row of 16)
  v += dpp::move<0x112, 0xF, 0xF>(v);  // row_shr:2
  v += dpp::move<0x114, 0xF, 0xF>(v);  // row_shr:4
  v += dpp::move<0x118, 0xF, 0xF>(v);  // row_shr:8: lane 15 of each row holds the row sum
  v += dpp::move<0x142, 0xA, 0xF>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp::move<0x143, 0xC, 0xF>(v);  // row_bcast:31 into rows 2 and 3: lane 63 holds the total
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

}  // namespace bd
