// plate_layout.hpp -- where the 13 values of a flat-plate node-pair block sit
// in the CSR rows of its owned node, and the wide stores that write them.
//
// The six rows of a node a with nb node-pair blocks hold 13 nb + 1 values from
// the CSR position of row 6a, block k (neighbours ascending) at
//   row X:          2k, 2k + 1              (columns X, Y of the neighbour)
//   row Y:          2 nb + 2k, + 1
//   row Z:          4 nb + 3k, + 1, + 2     (columns Z, ThX, ThY)
//   row ThX:        7 nb + 3k, + 1, + 2
//   row ThY:       10 nb + 3k, + 1, + 2
//   row ThZ:       13 nb                   (the drilling diagonal, one value)
// Adjacent values go out together: a 16-byte store per X / Y row pair and a
// 16-byte plus an 8-byte store per Z / ThX / ThY triple, so a block takes 8
// stores instead of 13 and each memory request carries more of a line.
#pragma once

#include <hip/hip_runtime.h>

namespace femdev {

typedef unsigned int plate_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int plate_u32x4 __attribute__((ext_vector_type(4)));

__device__ inline void plate_store2(__amdgpu_buffer_rsrc_t rsrc, uint32_t byte_off, double a, double b) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(plate_u32x4, make_double2(a, b)), rsrc, byte_off, 0, 0);
}

__device__ inline void plate_store1(__amdgpu_buffer_rsrc_t rsrc, uint32_t byte_off, double a) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(plate_u32x2, a), rsrc, byte_off, 0, 0);
}

// block k of a node with nb blocks; base: byte offset of the node's row 6a
__device__ inline void plate_store_block(__amdgpu_buffer_rsrc_t rsrc, uint32_t base, int k, int nb, const double o[13]) {
  const uint32_t b = base + 8u * uint32_t(2 * k);
  const uint32_t t = base + 8u * uint32_t(4 * nb + 3 * k);
  const uint32_t row3 = 8u * uint32_t(3 * nb);  // between the Z, ThX and ThY rows
  plate_store2(rsrc, b, o[0], o[1]);
  plate_store2(rsrc, b + 8u * uint32_t(2 * nb), o[2], o[3]);
  plate_store2(rsrc, t, o[4], o[5]);
  plate_store1(rsrc, t + 16u, o[6]);
  plate_store2(rsrc, t + row3, o[7], o[8]);
  plate_store1(rsrc, t + row3 + 16u, o[9]);
  plate_store2(rsrc, t + 2u * row3, o[10], o[11]);
  plate_store1(rsrc, t + 2u * row3 + 16u, o[12]);
}

}  // namespace femdev
