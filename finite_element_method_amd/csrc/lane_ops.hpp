// lane_ops.hpp -- exact moves of doubles between the lanes of a quad.
#pragma once

#include <hip/hip_runtime.h>

namespace femdev {

// every lane of a quad (lanes 4j .. 4j + 3) receives lane L's value: a DPP
// quad broadcast of both 32-bit halves, so the bits move unchanged
template <int L>
__device__ inline double quad_bcast(double v) {
  typedef unsigned int qb_u32x2 __attribute__((ext_vector_type(2)));
  const qb_u32x2 x = __builtin_bit_cast(qb_u32x2, v);
  qb_u32x2 r;
  r.x = uint32_t(__builtin_amdgcn_mov_dpp(int(x.x), L * 0x55, 0xF, 0xF, false));
  r.y = uint32_t(__builtin_amdgcn_mov_dpp(int(x.y), L * 0x55, 0xF, 0xF, false));
  return __builtin_bit_cast(double, r);
}

}  // namespace femdev
