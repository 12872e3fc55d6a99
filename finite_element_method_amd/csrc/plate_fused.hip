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

constexpr int PF_SLOT = 52;

__global__ __launch_bounds__(PF_MAX_PAIRS) void plate_fused_kernel(
    const double* __restrict__ xyz, const uint32_t* __restrict__ conn, const double* __restrict__ params,
    const uint32_t* __restrict__ node_ptr, const uint32_t* __restrict__ own_node,
    const uint32_t* __restrict__ own_info, const uint32_t* __restrict__ blk_ptr, const uint32_t* __restrict__ blk,
    const uint32_t* __restrict__ pair_ptr, const uint32_t* __restrict__ pairs, const int64_t* __restrict__ row_ptr,
    double* __restrict__ val) {
  __shared__ double stage[PF_MAX_PAIRS * PF_SLOT];
  __shared__ int stage_p[PF_MAX_PAIRS];
  const int patch = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t pb = pair_ptr[patch], npairs = pair_ptr[patch + 1] - pb;

  // ---- phase 1: row blocks of (owned node, element) pairs
  if (t < int(npairs)) {
    const uint32_t ref = pairs[pb + t];
    const uint32_t e = ref & 0x3FFFFFFFu;
    const int p = int(ref >> 30);
    const uint4 c = *reinterpret_cast<const uint4*>(conn + 4 * size_t(e));
    const double x3 = xyz[3 * size_t(c.z)], y3 = xyz[3 * size_t(c.z) + 1];
    double xl[6];
    xl[0] = xyz[3 * size_t(c.x)] - x3; xl[1] = xyz[3 * size_t(c.x) + 1] - y3;
    xl[2] = xyz[3 * size_t(c.y)] - x3; xl[3] = xyz[3 * size_t(c.y) + 1] - y3;
    xl[4] = xyz[3 * size_t(c.w)] - x3; xl[5] = xyz[3 * size_t(c.w) + 1] - y3;
    const double4 prm = *reinterpret_cast<const double4*>(params + 4 * size_t(e));
    const fem::PlateMat mat = fem::plate_material(prm.x, prm.y, prm.z, prm.w);
    double acc[PF_SLOT];
#pragma unroll
    for (int k = 0; k < PF_SLOT; ++k) acc[k] = 0.0;
    const double g = fem::plate_gauss();
#pragma unroll 1
    for (int ip = 0; ip < 4; ++ip) {
      const double r = (ip == 0 || ip == 3) ? g : -g;
      const double s = (ip < 2) ? g : -g;
      fem::PlateIP gi;
      fem::plate_ip(xl, r, s, gi);
      fem::plate_row_add_ip(gi, mat, p, acc);
    }
    double* out = stage + t * PF_SLOT;
#pragma unroll
    for (int k = 0; k < PF_SLOT; ++k) out[k] = acc[k];
    stage_p[t] = p;
  }
  __syncthreads();

  // ---- phase 2: owned CSR segments
  const int lane = t & 63, wave = t >> 6, nwaves = blockDim.x >> 6;
  const uint32_t n0 = node_ptr[patch], n1 = node_ptr[patch + 1];
  for (uint32_t o = n0 + wave; o < n1; o += nwaves) {
    const uint32_t a = own_node[o], info = own_info[o];
    const int deg = int(info & 255u), nb = int((info >> 8) & 255u), kself = int((info >> 16) & 255u);
    const int sb = int(info >> 24);
    const int seg = 13 * nb + 1;
    const int64_t base = row_ptr[6 * int64_t(a)];
    const uint32_t* bl = blk + blk_ptr[o];
    for (int pos = lane; pos < seg; pos += 64) {
      double acc;
      if (pos == seg - 1) {
        // drilling row: KROT6 of every element, in insertion order
        acc = fem::KROT6;
        for (int r = 1; r < deg; ++r) acc += fem::KROT6;
      } else {
        int k, idx;
        if (pos < 4 * nb) {
          const int i = pos >= 2 * nb;
          const int rem = pos - 2 * nb * i;
          k = rem >> 1;
          idx = 2 * i + (rem & 1);
        } else {
          const int rem = pos - 4 * nb;
          const int i = rem / (3 * nb);
          const int rr = rem - i * 3 * nb;
          k = rr / 3;
          idx = 4 + 3 * i + (rr - 3 * k);
        }
        if (k == kself) {
          acc = stage[sb * PF_SLOT + 13 * stage_p[sb] + idx];
          for (int r = 1; r < deg; ++r) acc += stage[(sb + r) * PF_SLOT + 13 * stage_p[sb + r] + idx];
        } else {
          const uint32_t d = bl[k];
          const int cnt = int(d & 7u);
          int r = int((d >> 3) & 7u), q = int((d >> 6) & 3u);
          acc = stage[(sb + r) * PF_SLOT + 13 * q + idx];
          for (int m = 1; m < cnt; ++m) {
            r = int((d >> (3 + 5 * m)) & 7u);
            q = int((d >> (6 + 5 * m)) & 3u);
            acc += stage[(sb + r) * PF_SLOT + 13 * q + idx];
          }
        }
      }
      val[base + pos] = acc;
    }
  }
}

std::string upload_plate_maps(Device* d, const PlateMaps& pm) {
  d->n_patches = pm.n_patches;
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
    plate_fused_kernel<<<d->n_patches, PF_MAX_PAIRS, 0, s>>>(
        d->xyz, d->conn[PLATE], d->params[PLATE], d->pf_node_ptr, d->pf_node, d->pf_info, d->pf_blk_ptr, d->pf_blk,
        d->pf_pair_ptr, d->pf_pairs, d->row_ptr, val);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(d->ev1, s));
  d->timed = true;
  return "";
}

}  // namespace femdev
