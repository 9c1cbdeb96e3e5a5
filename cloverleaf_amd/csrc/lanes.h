// Lane shifts across a wave64 for the marching kernels (kernels_fused.hip,
// kernels_entry.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace clv {

// Lane shifts across the whole wave64 with DPP (GFX9 wave_shr:1 / wave_shl:1):
// one VALU move per 32-bit half instead of an LDS-crossbar bpermute.  The lane
// shifted in at the wave's edge keeps its own value; edge lanes never produce
// owned results, so that value is never used.
__device__ __forceinline__ double dpp_shift(double v, int ctrl_is_shr) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  if (ctrl_is_shr) {
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);  // wave_shr:1
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);  // wave_shl:1
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
  }
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double up1(double v) { return dpp_shift(v, 1); }          // value of lane - 1
__device__ __forceinline__ double up2(double v) { return dpp_shift(dpp_shift(v, 1), 1); }  // value of lane - 2
__device__ __forceinline__ double dn1(double v) { return dpp_shift(v, 0); }          // value of lane + 1
__device__ __forceinline__ double dn2(double v) { return dpp_shift(dpp_shift(v, 0), 0); }  // value of lane + 2

}  // namespace clv
