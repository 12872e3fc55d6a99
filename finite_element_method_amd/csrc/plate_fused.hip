// plate_fused.hip -- fused assembly of flat Mindlin/MITC4 plate meshes.
//
// One workgroup per patch of <= 32 owned nodes (a spatially compact group:
// nodes in Morton order, so each patch covers a small 2D tile of the mesh).
// Let (a, r) be "owned node a and its r-th element in insertion order".
//   phase 1: one lane per pair (a, r) computes the row block of node a in
//            that element's stiffness (all four column nodes, 52 values: the
//            membrane, bending and MITC4 shear terms summed over the 2x2 Gauss
//            points, plate.rs:513-724) and stages it in LDS.  Each (element,
//            node) row block is computed exactly once over the whole grid.
//   phase 2: each owned node's CSR segment (rows 6a..6a+5, 13 nb + 1 values)
//            is summed from the staged blocks of its elements in insertion
//            order and written once, wave-contiguous.  No atomics: the result
//            is bitwise reproducible.
// The frame of every element is exactly the identity (checked when the maps
// are built), so local and global frames coincide and the drilling entry of
// every owned node is the sum of KROT6 over its elements.
#include <hip/hip_runtime.h>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

constexpr int PF_BS = 14;            // doubles per staged block (13 used; 16-byte aligned)
constexpr int PF_SLOT = 4 * PF_BS;    // doubles per (owned node, element) row block

// PAIRS = threads per workgroup = max (owned node, element) pairs of a patch;
// NODES = max owned nodes; BLKS = max node-pair blocks of a patch.
template <int PAIRS, int NODES, int BLKS>
__global__ __launch_bounds__(PAIRS) void plate_fused_kernel(
    const double* __restrict__ xyz, const uint32_t* __restrict__ conn, const double* __restrict__ params,
    const uint32_t* __restrict__ node_ptr, const uint32_t* __restrict__ own_node,
    const uint32_t* __restrict__ own_info, const uint32_t* __restrict__ blk_ptr, const uint32_t* __restrict__ blk,
    const uint32_t* __restrict__ pair_ptr, const uint32_t* __restrict__ pairs, const int64_t* __restrict__ row_ptr,
    double* __restrict__ val) {
  __shared__ double2 stage2[PAIRS * PF_SLOT / 2];
  __shared__ int stage_p[PAIRS];
  __shared__ int64_t s_base[NODES];
  __shared__ uint32_t s_info[NODES];
  __shared__ uint32_t s_boff[NODES];
  double* stage = reinterpret_cast<double*>(stage2);
  const int patch = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t pb = pair_ptr[patch], npairs = pair_ptr[patch + 1] - pb;
  const uint32_t n0 = node_ptr[patch], nown = node_ptr[patch + 1] - n0;
  const uint32_t b0 = blk_ptr[n0], nblk = blk_ptr[n0 + nown] - b0;

  // ---- metadata of the owned nodes
  if (t < int(nown)) {
    const uint32_t a = own_node[n0 + t];
    s_base[t] = row_ptr[6 * int64_t(a)];
    s_info[t] = own_info[n0 + t];
    s_boff[t] = blk_ptr[n0 + t] - b0;
  }

  // ---- phase 1: row blocks of (owned node, element) pairs
  if (t < int(npairs)) {
    const uint32_t ref = pairs[pb + t];
    const uint32_t e = ref & 0x3FFFFFFFu;
    const int p = int(ref >> 30);
    const uint4 c = *reinterpret_cast<const uint4*>(conn + 4 * size_t(e));
    const double4 prm = *reinterpret_cast<const double4*>(params + 4 * size_t(e));
    const double x3 = xyz[3 * size_t(c.z)], y3 = xyz[3 * size_t(c.z) + 1];
    double xl[6];
    xl[0] = xyz[3 * size_t(c.x)] - x3; xl[1] = xyz[3 * size_t(c.x) + 1] - y3;
    xl[2] = xyz[3 * size_t(c.y)] - x3; xl[3] = xyz[3 * size_t(c.y) + 1] - y3;
    xl[4] = xyz[3 * size_t(c.w)] - x3; xl[5] = xyz[3 * size_t(c.w) + 1] - y3;
    const fem::PlateMat mat = fem::plate_material(prm.x, prm.y, prm.z, prm.w);
    double row[PF_SLOT];
    fem::plate_row_folded<PF_BS>(xl, mat, p, row);
#pragma unroll
    for (int k = 0; k < PF_SLOT / 2; ++k) stage2[t * (PF_SLOT / 2) + k] = make_double2(row[2 * k], row[2 * k + 1]);
    stage_p[t] = p;
  }
  __syncthreads();

  // ---- phase 2: one lane per owned node-pair block; contributions summed in
  //      element-insertion order, 13 (14 on the diagonal) CSR values written once
  for (uint32_t bi = t; bi < nblk; bi += PAIRS) {
    const uint32_t d = blk[b0 + bi];
    const int n = int((d >> 23) & 63u);
    const uint32_t info = s_info[n];
    const int deg = int(info & 255u), nb = int((info >> 8) & 255u), sb = int(info >> 24);
    const int k = int(bi - s_boff[n]);
    const bool self = (d >> 29) & 1u;
    const int cnt = self ? deg : int(d & 7u);
    double acc[14];
#pragma unroll
    for (int m = 0; m < 14; ++m) acc[m] = 0.0;
    for (int m = 0; m < cnt; ++m) {
      int r, q;
      if (self) { r = m; q = stage_p[sb + m]; }
      else { r = int((d >> (3 + 5 * m)) & 7u); q = int((d >> (6 + 5 * m)) & 3u); }
      const double2* src = stage2 + ((sb + r) * PF_SLOT + PF_BS * q) / 2;
#pragma unroll
      for (int h = 0; h < 7; ++h) {
        const double2 v = src[h];
        if (m == 0) { acc[2 * h] = v.x; acc[2 * h + 1] = v.y; }
        else { acc[2 * h] += v.x; acc[2 * h + 1] += v.y; }
      }
    }
    double* o = val + s_base[n];
    o[2 * k] = acc[0];
    o[2 * k + 1] = acc[1];
    o[2 * nb + 2 * k] = acc[2];
    o[2 * nb + 2 * k + 1] = acc[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      o[4 * nb + 3 * k + j] = acc[4 + j];
      o[7 * nb + 3 * k + j] = acc[7 + j];
      o[10 * nb + 3 * k + j] = acc[10 + j];
    }
    if (self) {
      // drilling row: KROT6 of every element, in insertion order
      double dr = fem::KROT6;
      for (int r = 1; r < deg; ++r) dr += fem::KROT6;
      o[13 * nb] = dr;
    }
  }
}

std::string upload_plate_maps(Device* d, const PlateMaps& pm) {
  d->n_patches = pm.n_patches;
  d->pf_pairs_per_patch = pm.pairs_per_patch;
  std::string e;
  if (!(e = dupload(d->pf_node_ptr, pm.patch_node_ptr.data(), pm.patch_node_ptr.size())).empty()) return e;
  if (!(e = dupload(d->pf_node, pm.own_node.data(), pm.own_node.size())).empty()) return e;
  if (!(e = dupload(d->pf_info, pm.own_info.data(), pm.own_info.size())).empty()) return e;
  if (!(e = dupload(d->pf_blk_ptr, pm.own_blk_ptr.data(), pm.own_blk_ptr.size())).empty()) return e;
  if (!(e = dupload(d->pf_blk, pm.blk.data(), pm.blk.size())).empty()) return e;
  if (!(e = dupload(d->pf_pair_ptr, pm.patch_pair_ptr.data(), pm.patch_pair_ptr.size())).empty()) return e;
  return dupload(d->pf_pairs, pm.pairs.data(), pm.pairs.size());
}

std::string assemble_plate_fused(Device* d, double* val, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->n_patches > 0) {
#define PF_LAUNCH(P, N)                                                                                        \
  plate_fused_kernel<P, N, N * 16><<<d->n_patches, P, 0, s>>>(                                                 \
      d->xyz, d->conn[PLATE], d->params[PLATE], d->pf_node_ptr, d->pf_node, d->pf_info, d->pf_blk_ptr,        \
      d->pf_blk, d->pf_pair_ptr, d->pf_pairs, d->row_ptr, val)
    if (d->pf_pairs_per_patch == 64) PF_LAUNCH(64, 16);
    else if (d->pf_pairs_per_patch == 128) PF_LAUNCH(128, 32);
    else PF_LAUNCH(256, 64);
#undef PF_LAUNCH
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(d->ev1, s));
  d->timed = true;
  return "";
}

}  // namespace femdev
