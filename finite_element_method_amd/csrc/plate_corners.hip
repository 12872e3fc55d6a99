// plate_corners.hip -- fused assembly of flat Mindlin/MITC4 plate meshes, one
// lane per (owned node, incident element) corner.
//
// A workgroup of 256 threads takes a patch: up to 64 owned nodes (Morton
// order), four lanes per node (its elements in insertion order, idle lanes
// padding nodes with fewer elements) and the <= 64 elements touching them.
//   step 1: lanes (element, Gauss point) compute the point kinematics
//           (det J, dh/dx, dh/dy of the four nodes, the two values each of
//           the MITC4 multipliers) and the element constants -> LDS
//   step 2: lane (a, e) with a at corner p of e folds the 2 x 2 Gauss rule
//           into the six integrals of the element's node pairs (p, q) for
//           all four q (plate.rs:513-724) and forms the four 13-entry blocks
//           of row p.  Blocks with several contributing elements are summed
//           across the node's lanes in element-insertion order:
//             diagonal block (a, a): lane 0 adds the self blocks of lanes
//               1, 2, 3 in turn (DPP quad broadcasts);
//             a block shared by two elements: one lane exports it, the other
//               pulls it (one ds_bpermute round) and adds it (two terms: the
//               sum is the same in either order);
//           every other block has one contributor and is stored as is.
// Each CSR value is written once, by the lane that holds its sum: no atomics,
// no staging.  The arithmetic per value is the same as in plate_blocks.hip,
// term for term, so the two kernels give identical values.
#include <hip/hip_runtime.h>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"
#include "lane_ops.hpp"
#include "plate_layout.hpp"

namespace femdev {

namespace {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int PC_THREADS = 4 * PC_NODES;
// LDS per element, in double2: per Gauss point (w, a+), (a-, b+), (b-, 0),
// then (dx, dy) of nodes 0..3 (7 each, 28), the material (4), the shear-vector
// pairs e12, e43, e14, e23 (4), one pad: 37 = an odd number of 16-byte words,
// so that lanes reading different elements spread over the banks.
constexpr int PC_IP = 7;
constexpr int PC_MAT = 4 * PC_IP;
constexpr int PC_SHEAR = PC_MAT + 4;
constexpr int PC_EL = PC_SHEAR + 5;
static_assert(PC_THREADS == 4 * PC_MAX_ELEMS, "one lane per (element, point)");
static_assert(PC_MAX_ELEMS * PC_EL * 16 <= 40 * 1024, "four workgroups per CU");

struct CornerArgs {
  const double* __restrict__ xyz;
  const double* __restrict__ params;
  const uint2* __restrict__ hdr;    // per patch: elem_off, n_elems
  const uint4* __restrict__ pconn;  // per patch element: node indices
  const uint32_t* __restrict__ peid;
  const uint4* __restrict__ lane;   // per patch and lane: see PlateCornerMaps
  int n_patches;
  double* __restrict__ val;
  uint32_t val_bytes;
};

__device__ inline double lane_pull(double v, int byte_addr) {
  const u32x2 x = __builtin_bit_cast(u32x2, v);
  u32x2 r;
  r.x = uint32_t(__builtin_amdgcn_ds_bpermute(byte_addr, int(x.x)));
  r.y = uint32_t(__builtin_amdgcn_ds_bpermute(byte_addr, int(x.y)));
  return __builtin_bit_cast(double, r);
}

// block of node pair (p, q) of the element at E from the six integrals of slot s
__device__ inline void slot_block(const double2* E, const fem::PlateMat& mat, const double vp[3], const double up[3],
                                  int q, double Mxx, double Mxy, double Myx, double Myy, double Aa, double Bb,
                                  double o[13]) {
  const double2 ve = E[PC_SHEAR + (q < 2 ? 0 : 1)];
  const double2 ue = E[PC_SHEAR + ((q == 0 || q == 3) ? 2 : 3)];
  const double vq[3] = {(q == 0 || q == 3) ? 0.5 : -0.5, ve.x, ve.y};
  const double uq[3] = {q < 2 ? 0.5 : -0.5, ue.x, ue.y};
  fem::plate_block_entries(Mxx, Mxy, Myx, Myy, Aa, Bb, mat, vp, up, vq, uq, o);
}

__global__ __launch_bounds__(PC_THREADS, 4) void plate_corners_kernel(CornerArgs A) {
  __shared__ double2 L[PC_MAX_ELEMS * PC_EL];
  const int t = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A.val, 0, int(A.val_bytes), 0x00020000);
  for (int patch = blockIdx.x; patch < A.n_patches; patch += gridDim.x) {
    const uint2 h = A.hdr[patch];
    const uint32_t elem_off = h.x, nel = h.y;
    const uint4 d = A.lane[size_t(patch) * PC_THREADS + t];
    __syncthreads();  // the previous patch's LDS reads are complete

    // ---- step 1: point kinematics and element constants
    {
      const int le = t >> 2, ip = t & 3;
      if (le < int(nel)) {
        const uint4 c = A.pconn[elem_off + le];
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
        double2* go = L + le * PC_EL + ip * PC_IP;
        go[0] = make_double2(g.w, g.al[0]);
        go[1] = make_double2(g.al[2], g.be[0]);
        go[2] = make_double2(g.be[1], 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) go[3 + a] = make_double2(g.dx[a], g.dy[a]);
        double2* ed = L + le * PC_EL;
        if (ip == 0) {
          const uint32_t eid = A.peid[elem_off + le];
          const double4 prm = *reinterpret_cast<const double4*>(A.params + 4 * size_t(eid));
          const fem::PlateMat m = fem::plate_material(prm.x, prm.y, prm.z, prm.w);
          ed[PC_MAT] = make_double2(m.cm00, m.cm01);
          ed[PC_MAT + 1] = make_double2(m.cm22, m.cb00);
          ed[PC_MAT + 2] = make_double2(m.cb01, m.cb22);
          ed[PC_MAT + 3] = make_double2(m.cs, 0.0);
        } else if (ip == 1) {
          fem::PlateShear sv;
          fem::plate_shear_vectors(xl, sv);
          ed[PC_SHEAR] = make_double2(sv.v[0][1], sv.v[0][2]);      // e12
          ed[PC_SHEAR + 1] = make_double2(sv.v[2][1], sv.v[2][2]);  // e43
          ed[PC_SHEAR + 2] = make_double2(sv.u[0][1], sv.u[0][2]);  // e14
          ed[PC_SHEAR + 3] = make_double2(sv.u[1][1], sv.u[1][2]);  // e23
        }
      }
    }
    __syncthreads();

    // ---- step 2: the four blocks of row p of element e at node a
    // (all lanes run it, so that the cross-lane exchanges see every lane;
    // idle lanes read element 0 and store nothing)
    const uint32_t y = d.y, z = d.z;
    const bool active = (y & 127u) != 127u;
    const double2* E = L + (active ? int(y & 127u) : 0) * PC_EL;
    const int p = int((y >> 7) & 3u);
    const int qs[4] = {p, int((y >> 9) & 3u), int((y >> 11) & 3u), int((y >> 13) & 3u)};
    const int nb = int((y >> 15) & 31u);
    // alpha(q) = (1 + s) mr for q = 0, 1 and (1 - s) mr for q = 2, 3;
    // beta(q) = (1 + r) ms for q = 0, 3 and (1 - r) ms for q = 1, 2
    bool aplus[4], bplus[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      aplus[s] = qs[s] < 2;
      bplus[s] = ((9u >> qs[s]) & 1u) != 0u;
    }
    double Mxx[4], Mxy[4], Myx[4], Myy[4], Aa[4], Bb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) Mxx[s] = Mxy[s] = Myx[s] = Myy[s] = Aa[s] = Bb[s] = 0.0;
    // one point at a time: keeps the loads of a single point in flight, not
    // all four (the accumulators already hold 48 registers)
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
      const double2* G = E + i * PC_IP;
      const double2 s0 = G[0], s1 = G[1], s2 = G[2];
      const double w = s0.x;
      double2 dq[4];
      double al[4], be[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        dq[s] = G[3 + qs[s]];
        al[s] = aplus[s] ? s0.y : s1.x;
        be[s] = bplus[s] ? s1.y : s2.x;
      }
      const double wdx = w * dq[0].x, wdy = w * dq[0].y, wal = w * al[0], wbe = w * be[0];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        Mxx[s] = __builtin_fma(wdx, dq[s].x, Mxx[s]);
        Mxy[s] = __builtin_fma(wdx, dq[s].y, Mxy[s]);
        Myx[s] = __builtin_fma(wdy, dq[s].x, Myx[s]);
        Myy[s] = __builtin_fma(wdy, dq[s].y, Myy[s]);
        Aa[s] = __builtin_fma(wal, al[s], Aa[s]);
        Bb[s] = __builtin_fma(wbe, be[s], Bb[s]);
      }
    }
    fem::PlateMat mat;
    {
      const double2 e0 = E[PC_MAT], e1 = E[PC_MAT + 1], e2 = E[PC_MAT + 2], e3 = E[PC_MAT + 3];
      mat.cm00 = e0.x; mat.cm01 = e0.y; mat.cm22 = e1.x; mat.cb00 = e1.y; mat.cb01 = e2.x; mat.cb22 = e2.y;
      mat.cs = e3.x;
    }
    double vp[3], up[3];
    {
      const double2 ve = E[PC_SHEAR + (p < 2 ? 0 : 1)];
      const double2 ue = E[PC_SHEAR + ((p == 0 || p == 3) ? 2 : 3)];
      vp[0] = mat.cs * ((p == 0 || p == 3) ? 0.5 : -0.5);
      vp[1] = mat.cs * ve.x;
      vp[2] = mat.cs * ve.y;
      up[0] = mat.cs * (p < 2 ? 0.5 : -0.5);
      up[1] = mat.cs * ue.x;
      up[2] = mat.cs * ue.y;
    }
    const uint32_t base = d.x * 8u;  // 32-bit byte offsets (checked when the maps are built)

    // slots are finished in the order 3, (1, 2), 0 so that the integrals of
    // the later slots are the only values that stay live
    __builtin_amdgcn_sched_barrier(0);
    // slot 3: a block with this element as its only contributor
    {
      double o3[13];
      slot_block(E, mat, vp, up, qs[3], Mxx[3], Mxy[3], Myx[3], Myy[3], Aa[3], Bb[3], o3);
      if (active) plate_store_block(rsrc, base, int((z >> 15) & 31u), nb, o3);
    }
    __builtin_amdgcn_sched_barrier(0);
    // slots 1 and 2: a block exported to another lane (or kept) and a block
    // completed with the pulled one (or kept)
    {
      double o1[13], o2[13];
      slot_block(E, mat, vp, up, qs[1], Mxx[1], Mxy[1], Myx[1], Myy[1], Aa[1], Bb[1], o1);
      slot_block(E, mat, vp, up, qs[2], Mxx[2], Mxy[2], Myx[2], Myy[2], Aa[2], Bb[2], o2);
      const int src = ((t & 63) & ~3) | int((y >> 26) & 3u);
      const bool own = (y >> 25) & 1u;
#pragma unroll
      for (int i = 0; i < 13; ++i) {
        const double x = lane_pull(o1[i], 4 * src);
        o2[i] = own ? o2[i] + x : o2[i];
      }
      if (active) {
        if ((y >> 24) & 1u) plate_store_block(rsrc, base, int((z >> 5) & 31u), nb, o1);
        plate_store_block(rsrc, base, int((z >> 10) & 31u), nb, o2);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // diagonal block: lane 0 of the quad sums the self blocks of lanes 0 .. deg - 1
    {
      double o[13];
      slot_block(E, mat, vp, up, qs[0], Mxx[0], Mxy[0], Myx[0], Myy[0], Aa[0], Bb[0], o);
      const int deg = int((y >> 20) & 7u);  // 0 on every lane but the node's lane 0
#pragma unroll
      for (int i = 0; i < 13; ++i) {
        // all lanes of the wave read o[i] before any lane replaces it
        const double x1 = quad_bcast<1>(o[i]), x2 = quad_bcast<2>(o[i]), x3 = quad_bcast<3>(o[i]);
        double v = o[i];
        v = deg > 1 ? v + x1 : v;
        v = deg > 2 ? v + x2 : v;
        v = deg > 3 ? v + x3 : v;
        o[i] = v;
      }
      if (active && ((y >> 23) & 1u)) {
        plate_store_block(rsrc, base, int(z & 31u), nb, o);
        double dr = fem::KROT6;
        for (int r = 1; r < deg; ++r) dr += fem::KROT6;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, dr), rsrc, base + 8u * uint32_t(13 * nb), 0,
                                              0);
      }
    }
  }
}

}  // namespace

std::string upload_plate_corner_maps(Device* d, const PlateCornerMaps& pm, int wg_per_cu) {
  d->pb_patches = pm.n_patches;
  std::string e;
  if (!(e = dupload(d->pb_hdr, pm.hdr.data(), pm.hdr.size())).empty()) return e;
  if (!(e = dupload(d->pb_conn, pm.pconn.data(), pm.pconn.size())).empty()) return e;
  if (!(e = dupload(d->pb_eid, pm.peid.data(), pm.peid.size())).empty()) return e;
  if (!(e = dupload(d->pb_blk, pm.lane.data(), pm.lane.size())).empty()) return e;
  return set_plate_block_grid(d, wg_per_cu);
}

std::string assemble_plate_corners(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->pb_patches > 0) {
    CornerArgs a;
    a.xyz = d->xyz;
    a.params = d->params[PLATE];
    a.hdr = reinterpret_cast<const uint2*>(d->pb_hdr);
    a.pconn = reinterpret_cast<const uint4*>(d->pb_conn);
    a.peid = d->pb_eid;
    a.lane = reinterpret_cast<const uint4*>(d->pb_blk);
    a.n_patches = d->pb_patches;
    a.val = val;
    a.val_bytes = uint32_t(d->nnz * 8);
    plate_corners_kernel<<<d->pb_grid, PC_THREADS, 0, s>>>(a);
    HIP_TRY(hipGetLastError());
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

}  // namespace femdev
