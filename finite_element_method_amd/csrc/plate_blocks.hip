// plate_blocks.hip -- fused assembly of flat Mindlin/MITC4 plate meshes, one
// lane per output node-pair block.
//
// A workgroup of 256 threads takes a patch (owned nodes in Morton order, up
// to 256 node-pair blocks, the <= PB_ELEMS elements touching them):
//   step 1: lanes (element, Gauss point) compute the point kinematics (det J,
//           dh/dx, dh/dy, MITC4 multipliers) and the element constants
//           (constitutive scalars, shear vectors) of every patch element -> LDS
//   step 2: one lane per owned node-pair block (a, b) walks the elements that
//           couple a and b in insertion order (plate.rs:513-724 folded over
//           the four points), sums their 13-entry blocks and stores them into
//           the CSR rows of a (8 wide stores per block, plate_layout.hpp).  No atomics and no staging: every value is
//           written once, by the lane that summed it.
// Every element frame is exactly the identity (checked when the maps are
// built); the drilling entry of an owned node is the sum of KROT6 over its
// elements.  Operation order per value is the same as in plate_fused.hip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"
#include "plate_layout.hpp"

namespace femdev {

namespace {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int PB_THREADS = 256;
// LDS strides are odd multiples of 16 bytes, so that the 16-byte reads of
// lanes working on different elements spread over the banks.
constexpr int PB_GEO = 18;                 // doubles per (element, point): w, pad, then (dx, dy, al, be) of nodes 0..3
constexpr int PB_GEO_EL = 4 * PB_GEO + 2;  // doubles per element of point data
constexpr int PB_EDAT = 34;                // doubles per element: 7 constants, pad, (v[3], u[3]) of nodes 0..3, pad
static_assert(PB_THREADS >= 4 * PB_MAX_ELEMS, "one lane per (element, point)");
static_assert(PB_THREADS == PB_MAX_BLOCKS, "one lane per block");

struct BlockArgs {
  const double* __restrict__ xyz;
  const double* __restrict__ params;
  const uint4* __restrict__ hdr;    // per patch: elem_off, blk_off, n_elems, n_blocks
  const uint4* __restrict__ pconn;  // per patch element: node indices
  const uint32_t* __restrict__ peid;
  const uint4* __restrict__ blk;    // per block: see PlateBlockMaps
  int n_patches;
  double* __restrict__ val;
  uint32_t val_bytes;
};

struct BlockLds {
  double2 geo[PB_MAX_ELEMS * PB_GEO_EL / 2];
  double2 edat[PB_MAX_ELEMS * PB_EDAT / 2];
};
static_assert(sizeof(BlockLds) <= 40 * 1024, "four workgroups per CU");

__global__ __launch_bounds__(PB_THREADS, 4) void plate_blocks_kernel(BlockArgs A) {
  __shared__ BlockLds L;
  const int t = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A.val, 0, int(A.val_bytes), 0x00020000);
  for (int patch = blockIdx.x; patch < A.n_patches; patch += gridDim.x) {
    const uint4 h = A.hdr[patch];
    const uint32_t elem_off = h.x, blk_off = h.y, nel = h.z, nblk = h.w;
    const bool has_blk = uint32_t(t) < nblk;
    const uint4 desc = A.blk[blk_off + min(uint32_t(t), nblk - 1u)];
    __syncthreads();  // the previous patch's LDS reads are complete

    // ---- step 1: point kinematics and element constants
    const int le = t >> 2, ip = t & 3;
    if (le < int(nel)) {
      const uint4 c = A.pconn[elem_off + le];
      const uint32_t eid = A.peid[elem_off + le];
      const uint32_t n[4] = {c.x, c.y, c.z, c.w};
      double X[4][2];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double2 v = *reinterpret_cast<const double2*>(A.xyz + 3 * size_t(n[a]));
        X[a][0] = v.x;
        X[a][1] = v.y;
      }
      double xl[6];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int a = m < 2 ? m : 3;
        xl[2 * m] = X[a][0] - X[2][0];
        xl[2 * m + 1] = X[a][1] - X[2][1];
      }
      fem::PlatePoint g;
      fem::plate_point(xl, ip, g);
      double2* go = L.geo + (le * PB_GEO_EL + ip * PB_GEO) / 2;
      go[0] = make_double2(g.w, 0.0);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        go[1 + 2 * a] = make_double2(g.dx[a], g.dy[a]);
        go[2 + 2 * a] = make_double2(g.al[a], g.be[a]);
      }
      double2* ed = L.edat + le * (PB_EDAT / 2);
      if (ip == 0) {
        const double4 prm = *reinterpret_cast<const double4*>(A.params + 4 * size_t(eid));
        const fem::PlateMat m = fem::plate_material(prm.x, prm.y, prm.z, prm.w);
        ed[0] = make_double2(m.cm00, m.cm01);
        ed[1] = make_double2(m.cm22, m.cb00);
        ed[2] = make_double2(m.cb01, m.cb22);
        ed[3] = make_double2(m.cs, 0.0);
      } else {
        // lanes 1..3: the shear vectors of nodes ip - 1 (and 3 for lane 3)
        fem::PlateShear sv;
        fem::plate_shear_vectors(xl, sv);
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if (a == ip - 1 || (ip == 3 && a == 3)) {
            ed[4 + 3 * a] = make_double2(sv.v[a][0], sv.v[a][1]);
            ed[5 + 3 * a] = make_double2(sv.v[a][2], sv.u[a][0]);
            ed[6 + 3 * a] = make_double2(sv.u[a][1], sv.u[a][2]);
          }
      }
    }
    __syncthreads();

    // ---- step 2: one lane per node-pair block
    if (has_blk) {
      const uint32_t g0 = desc.x;  // CSR position of row 6a of the owned node
      const int k = int(desc.y & 31u), nb = int((desc.y >> 5) & 31u);
      const bool self = (desc.y >> 10) & 1u;
      const int cnt = int((desc.y >> 11) & 15u);
      double bs[13];
      for (int m = 0; m < cnt; ++m) {
        int cle, p, q;
        if (self) {
          const uint32_t cb = (m < 4 ? (desc.z >> (8 * m)) : (desc.w >> (8 * (m - 4)))) & 255u;
          cle = int(cb & 63u);
          p = q = int(cb >> 6);
        } else {
          const uint32_t cb = (m < 3 ? (desc.z >> (10 * m)) : desc.w) & 1023u;
          cle = int(cb & 63u);
          p = int((cb >> 6) & 3u);
          q = int(cb >> 8);
        }
        double Mxx = 0.0, Mxy = 0.0, Myx = 0.0, Myy = 0.0, Aa = 0.0, Bb = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double2* gi = L.geo + (cle * PB_GEO_EL + i * PB_GEO) / 2;
          const double w = gi[0].x;
          const double2 dp = gi[1 + 2 * p], ap = gi[2 + 2 * p];
          const double2 dq = gi[1 + 2 * q], aq = gi[2 + 2 * q];
          const double wdx = w * dp.x, wdy = w * dp.y, wal = w * ap.x, wbe = w * ap.y;
          Mxx = __builtin_fma(wdx, dq.x, Mxx);
          Mxy = __builtin_fma(wdx, dq.y, Mxy);
          Myx = __builtin_fma(wdy, dq.x, Myx);
          Myy = __builtin_fma(wdy, dq.y, Myy);
          Aa = __builtin_fma(wal, aq.x, Aa);
          Bb = __builtin_fma(wbe, aq.y, Bb);
        }
        const double2* ed = L.edat + cle * (PB_EDAT / 2);
        const double2 e0 = ed[0], e1 = ed[1], e2 = ed[2], e3 = ed[3];
        fem::PlateMat mat;
        mat.cm00 = e0.x; mat.cm01 = e0.y; mat.cm22 = e1.x; mat.cb00 = e1.y; mat.cb01 = e2.x; mat.cb22 = e2.y;
        mat.cs = e3.x;
        const double2 p0 = ed[4 + 3 * p], p1 = ed[5 + 3 * p], p2 = ed[6 + 3 * p];
        const double2 q0 = ed[4 + 3 * q], q1 = ed[5 + 3 * q], q2 = ed[6 + 3 * q];
        const double vp[3] = {mat.cs * p0.x, mat.cs * p0.y, mat.cs * p1.x};
        const double up[3] = {mat.cs * p1.y, mat.cs * p2.x, mat.cs * p2.y};
        const double vq[3] = {q0.x, q0.y, q1.x}, uq[3] = {q1.y, q2.x, q2.y};
        double o[13];
        fem::plate_block_entries(Mxx, Mxy, Myx, Myy, Aa, Bb, mat, vp, up, vq, uq, o);
#pragma unroll
        for (int i = 0; i < 13; ++i) bs[i] = (m == 0) ? o[i] : bs[i] + o[i];
      }
      const uint32_t base = g0 * 8u;  // 32-bit byte offsets (checked when the maps are built)
      plate_store_block(rsrc, base, k, nb, bs);
      if (self) {
        // drilling row: KROT6 of every element of the node, in insertion order
        double dr = fem::KROT6;
        for (int r = 1; r < cnt; ++r) dr += fem::KROT6;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, dr), rsrc, base + 8u * uint32_t(13 * nb), 0,
                                              0);
      }
    }
  }
}

}  // namespace

std::string upload_plate_block_maps(Device* d, const PlateBlockMaps& pm, int wg_per_cu) {
  d->pb_patches = pm.n_patches;
  std::string e;
  if (!(e = dupload(d->pb_hdr, pm.hdr.data(), pm.hdr.size())).empty()) return e;
  if (!(e = dupload(d->pb_conn, pm.pconn.data(), pm.pconn.size())).empty()) return e;
  if (!(e = dupload(d->pb_eid, pm.peid.data(), pm.peid.size())).empty()) return e;
  if (!(e = dupload(d->pb_blk, pm.blk.data(), pm.blk.size())).empty()) return e;
  return set_plate_block_grid(d, wg_per_cu);
}

std::string set_plate_block_grid(Device* d, int wg_per_cu) {
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->dev));
  d->pb_grid = wg_per_cu <= 0 ? d->pb_patches : std::max(1, std::min(d->pb_patches, cus * wg_per_cu));
  return "";
}

std::string assemble_plate_blocks(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->pb_patches > 0) {
    BlockArgs a;
    a.xyz = d->xyz;
    a.params = d->params[PLATE];
    a.hdr = reinterpret_cast<const uint4*>(d->pb_hdr);
    a.pconn = reinterpret_cast<const uint4*>(d->pb_conn);
    a.peid = d->pb_eid;
    a.blk = reinterpret_cast<const uint4*>(d->pb_blk);
    a.n_patches = d->pb_patches;
    a.val = val;
    a.val_bytes = uint32_t(d->nnz * 8);
    plate_blocks_kernel<<<d->pb_grid, PB_THREADS, 0, s>>>(a);
    HIP_TRY(hipGetLastError());
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

}  // namespace femdev
