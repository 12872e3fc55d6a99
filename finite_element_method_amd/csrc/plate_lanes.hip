// plate_lanes.hip -- fused assembly of flat Mindlin/MITC4 plate meshes whose
// nodes have at most four elements: one wave per patch of up to 16 owned
// nodes, balanced contribution lanes, values staged in LDS and written out as
// consecutive 16-byte chunks.
//
// Work of an owned node a: every node-pair block (a, b) is the sum, in element
// insertion order, of one 13-entry contribution per element holding a and b
// (plate.rs:513-724 folded over the 2x2 rule; the add_value order of
// methods_for_plate_data_handle.rs:134-212).  An interior node has 16
// contributions: its diagonal block 4, four edge neighbours 2 each, four
// corner neighbours 1 each.  The host packs each node's blocks into lanes of at
// most four contributions without splitting a block (4 | 2+2 | 2+2 | 1+1+1+1),
// so every lane runs the same four trips and no lane idles while another sums
// a long block.
//
// Steps of a wave:
//   1. element records: two lanes per patch element (integration points
//      {0, 1} and {2, 3}) form the 42 point integrals of every node-pair
//      bilinear form (sum_i w_i f_a f_b for f in dh/dx, dh/dy and the MITC4
//      multipliers, symmetric pairs once), combine their halves by a DPP lane
//      swap and store them with the constitutive constants and the shear
//      vectors in LDS;
//   2. trips: each lane forms its contributions from the records in registers
//      (13 entries per block from six integrals), sums each block and writes
//      the finished block into the node's staging segment in LDS;
//   3. copy-out: the node's CSR values (13 nb + 1, contiguous from row 6a) go
//      out as aligned 16-byte chunks from consecutive lanes, so every wave store
//      covers whole cache lines.
// No atomics; every value is written once.  Everything is deterministic and
// independent of how nodes are grouped into patches, so a mesh assembled
// alone and inside a batch gives the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

namespace {

typedef unsigned int pl_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int pl_u32x4 __attribute__((ext_vector_type(4)));

// element record (doubles): Mxx (10, p <= q), Myy (10), Mxy (16, 4 p + q),
// A (3, alpha classes), B (3, beta classes), shear edge vectors (8), material (cm00, cm01,
// cm22, t^2/12, cs), pad
constexpr int PL_REC = 58;
constexpr int R_MYY = 10, R_MXY = 20, R_A = 36, R_B = 39, R_SH = 42, R_MAT = 50;
constexpr int PL_NINT = 42;
constexpr int PL_STORE_NT = 2;  // buffer-store cache policy: non-temporal
constexpr int PL_GEO = 10;   // per wave element slot: x1, y1, x2, y2, x4, y4 about node 3; E, nu, t, k
constexpr int PL_GEO_U = 8;  // the same when E, nu, k are uniform: x1 .. y4; t; pad
static_assert(PL_REC % 2 == 0, "records start on 16-byte boundaries");
static_assert(2 * PL_ELEMS <= 64, "two lanes per element");
constexpr int PL_LDS = PL_ELEMS * PL_REC > PL_STAGE_VALUES ? PL_ELEMS * PL_REC : PL_STAGE_VALUES;
static_assert(PL_STAGE_VALUES % 2 == 0 && PL_STAGE_VALUES >= PL_NODES * (13 * PL_MAX_NB + 2), "staging");

// Everything a wave reads sits at fixed offsets from its index, so all its
// loads issue at once: no load waits for another's result.
struct LaneArgs {
  const double* __restrict__ geo;    // per wave PL_ELEMS slots of PL_GEO doubles
  // per wave PL_NODES: CSR position of row 6a of owned node s,
  // nb | (n_elems | n_nodes << 6 | trips << 11) << 4 | staging position << 18
  const uint2* __restrict__ nodes;
  const uint2* __restrict__ lanes;   // 64 lane descriptors per distinct set
  const uint32_t* __restrict__ lane_set;  // per wave: its descriptor set
  double* __restrict__ val;
  uint32_t val_bytes;
  uint32_t geo_stride;  // PL_GEO, or PL_GEO_U when E, nu, k are uniform (then in mat)
  double mat[3];
};

// inputs of wave patch w: the lane's descriptor, owned node (l mod 16) and
// the geometry slot of element l / 2
__device__ inline void load_patch(const LaneArgs& A, uint32_t w, int l, int le, uint2& desc, uint2& nd, double2& g0,
                                  double2& g1, double2& g2, double2& g3, double2& g4) {
  desc = A.lanes[size_t(A.lane_set[w]) * 64 + l];
  nd = A.nodes[size_t(w) * PL_NODES + (l & (PL_NODES - 1))];
  const double2* gs = reinterpret_cast<const double2*>(A.geo + (size_t(w) * PL_ELEMS + le) * A.geo_stride);
  g0 = gs[0];
  g1 = gs[1];
  g2 = gs[2];
  if (A.geo_stride == PL_GEO_U) {  // (E, nu) and (t, k) from the uniform values
    g3 = make_double2(A.mat[0], A.mat[1]);
    g4 = make_double2(gs[3].x, A.mat[2]);
  } else {
    g3 = gs[3];
    g4 = gs[4];
  }
}

// Fills the wave slots' geometry: local coordinates of nodes 1, 2, 4 about
// node 3 (the same differences the crate's plate frame takes, quad:340-378,
// with an identity frame) and the element parameters (t only when E, nu and k
// are uniform: stride PL_GEO_U).
__global__ __launch_bounds__(256) void plate_lanes_gather_kernel(const double* __restrict__ xyz,
                                                                 const double* __restrict__ params,
                                                                 const uint32_t* __restrict__ sconn,
                                                                 const uint32_t* __restrict__ seid, size_t n_slots,
                                                                 int stride, double* __restrict__ geo) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i >= n_slots) return;
  const uint32_t eid = seid[i];
  double* g = geo + i * stride;
  if (eid == 0xFFFFFFFFu) {  // unused slot
    for (int k = 0; k < stride; ++k) g[k] = 0.0;
    return;
  }
  const uint4 c = *reinterpret_cast<const uint4*>(sconn + 4 * i);
  const uint32_t n[4] = {c.x, c.y, c.z, c.w};
  double X[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    X[a][0] = xyz[3 * size_t(n[a])];
    X[a][1] = xyz[3 * size_t(n[a]) + 1];
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int a = m < 2 ? m : 3;
    g[2 * m] = X[a][0] - X[2][0];
    g[2 * m + 1] = X[a][1] - X[2][1];
  }
  if (stride == PL_GEO_U) {
    g[6] = params[4 * size_t(eid) + 2];
    g[7] = 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) g[6 + k] = params[4 * size_t(eid) + k];
  }
}

// index of the symmetric pair (a, b), a <= b, among the 10 pairs of 4 nodes
__device__ inline int sym_index(int a, int b) { return a * 4 - (a * (a - 1)) / 2 + (b - a); }

// both 32-bit halves of a double from the neighbouring lane (lane ^ 1), by DPP
__device__ inline double swap_pair(double v) {
  const pl_u32x2 x = __builtin_bit_cast(pl_u32x2, v);
  pl_u32x2 r;
  r.x = uint32_t(__builtin_amdgcn_mov_dpp(int(x.x), 0xB1, 0xF, 0xF, false));
  r.y = uint32_t(__builtin_amdgcn_mov_dpp(int(x.y), 0xB1, 0xF, 0xF, false));
  return __builtin_bit_cast(double, r);
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v with its sign flipped when neg is 1 (an exact sign change)
__device__ inline double with_sign(double v, int neg) {
  pl_u32x2 x = __builtin_bit_cast(pl_u32x2, v);
  x.y ^= uint32_t(neg) << 31;
  return __builtin_bit_cast(double, x);
}

// The 13 entries of the flat-plate block (p, q) (layout of plate_layout.hpp:
// (u,u) (u,v) | (v,u) (v,v) | (w,w) (w,tx) (w,ty) | (tx,.) | (ty,.)) from the
// six point integrals of the pair, plate.rs:513-672 folded over the 2x2 rule:
// membrane from cm * M, bending as t^2/12 times the matching membrane entry,
// MITC4 shear A cs v_p v_q^T + B cs u_p u_q^T with v_a = (sv_a / 2, ev_a),
// u_a = (su_a / 2, eu_a) (sv, su: sign bits).
__device__ inline void block_entries(double Mxx, double Mxy, double Myx, double Myy, double Acs, double Bcs,
                                     double cm00, double cm01, double cm22, double tb, double2 evp, double2 eup,
                                     double2 evq, double2 euq, int svp, int sup, int svq, int suq, double o[13]) {
  o[0] = __builtin_fma(cm00, Mxx, cm22 * Myy);
  o[1] = __builtin_fma(cm01, Mxy, cm22 * Myx);
  o[2] = __builtin_fma(cm01, Myx, cm22 * Mxy);
  o[3] = __builtin_fma(cm00, Myy, cm22 * Mxx);
  const double bxx = tb * o[3], bxy = -(tb * o[2]), byx = -(tb * o[1]), byy = tb * o[0];
  const double a[3] = {with_sign(0.5 * Acs, svp), Acs * evp.x, Acs * evp.y};
  const double b[3] = {with_sign(0.5 * Bcs, sup), Bcs * eup.x, Bcs * eup.y};
  const double vq[3] = {with_sign(0.5, svq), evq.x, evq.y};
  const double uq[3] = {with_sign(0.5, suq), euq.x, euq.y};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) o[4 + 3 * i + j] = __builtin_fma(a[i], vq[j], b[i] * uq[j]);
  o[8] += bxx;
  o[9] += bxy;
  o[11] += byx;
  o[12] += byy;
}

// Integration points 2h and 2h + 1 of lane half h, in the crate's order
// (g,g), (-g,g), (-g,-g), (g,-g): both have s = +g (h = 0) or -g (h = 1), and
// r = s, -s.  The quantities and formulas of fem::plate_point (det J,
// dh/dx, dh/dy quad:380-653, MITC4 multipliers plate.rs:424-447); the parts
// that depend on s only are formed once, and the square root of the gamma_rz
// length at r = -s comes from the other lane of the pair (whose s is -s).
__device__ inline void lane_points(const double xl[6], int half, fem::PlatePoint& P0, fem::PlatePoint& P1) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double gs = sqrt(double(fem::ONE_THIRD_F32));
  const double s = half ? -gs : gs;
  const double q4 = 0.25;
  const double x1q = x1 * q4, x2q = x2 * q4, x4q = x4 * q4, y1q = y1 * q4, y2q = y2 * q4, y4q = y4 * q4;
  const double j00 = ((((x1q + x1q * s) - x2q) - x2q * s) + x4q) - x4q * s;
  const double j01 = ((((y1q + y1q * s) - y2q) - y2q * s) + y4q) - y4q * s;
  const double dr[4] = {q4 + q4 * s, -(q4 + q4 * s), -(q4 - q4 * s), q4 - q4 * s};
  const double ax = x1 - x2 + x4, bx = x1 - x2 - x4, cx = x1 + x2 - x4;
  const double ay = y1 - y2 + y4, by = y1 - y2 - y4, cy = y1 + y2 - y4;
  const double sq_s = sqrt((ax + s * bx) * (ax + s * bx) + (ay + s * by) * (ay + s * by));
  const double r0 = s;
  const double sq_r0 = sqrt((cx + r0 * bx) * (cx + r0 * bx) + (cy + r0 * by) * (cy + r0 * by));
  const double sq_r1 = swap_pair(sq_r0);  // r = -s: the other lane's r0
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    fem::PlatePoint& o = j ? P1 : P0;
    const double r = j ? -s : s;
    const double j10 = ((((x1q + x1q * r) + x2q) - x2q * r) - x4q) - x4q * r;
    const double j11 = ((((y1q + y1q * r) + y2q) - y2q * r) - y4q) - y4q * r;
    const double det = j00 * j11 - j01 * j10;
    const double rdet = 1.0 / det;
    const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
    const double ds[4] = {q4 + q4 * r, q4 - q4 * r, -(q4 - q4 * r), -(q4 + q4 * r)};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      o.dx[a] = i00 * dr[a] + i01 * ds[a];
      o.dy[a] = i10 * dr[a] + i11 * ds[a];
    }
    const double r8 = 0.125 * rdet;
    const double mr = (j ? sq_r1 : sq_r0) * r8;
    const double ms = sq_s * r8;
    o.al[0] = (1.0 + s) * mr; o.al[1] = o.al[0]; o.al[2] = (1.0 - s) * mr; o.al[3] = o.al[2];
    o.be[0] = (1.0 + r) * ms; o.be[1] = (1.0 - r) * ms; o.be[2] = o.be[1]; o.be[3] = o.be[0];
    o.w = det;
  }
}

// block index k of item t of a lane descriptor (t < PL_TRIPS)
__device__ inline int item_block(uint64_t items, int t) { return int(uint32_t(items >> (13 * t + 9)) & 15u); }


__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void plate_lanes_kernel(
    LaneArgs A) {
  // element records during steps 1-2, the owned nodes' staged values after
  __shared__ double2 lds2[PL_LDS / 2];
  double* rec = reinterpret_cast<double*>(lds2);
  double* stage = rec;
  const double2* stage2 = lds2;
  const int l = threadIdx.x;
  const int le = l >> 1, half = l & 1;
  const uint32_t w = blockIdx.x;
  uint2 desc, nd;
  double2 g0, g1, g2, g3, g4;
  load_patch(A, w, l, le, desc, nd, g0, g1, g2, g3, g4);
  // nb | (n_elems | n_nodes << 6 | trips << 11) << 4 | staging position << 18
  const uint32_t h = uint32_t(__builtin_amdgcn_readfirstlane(int(nd.y))) >> 4;
  const int nel = int(h & 63u), nnod = int((h >> 6) & 31u), trips = int((h >> 11) & 7u);

  // ---- step 1: element records, two lanes per element
  {
    const bool live = le < nel;
    double I[PL_NINT];
#pragma unroll
    for (int k = 0; k < PL_NINT; ++k) I[k] = 0.0;
    const double xl[6] = {g0.x, g0.y, g1.x, g1.y, g2.x, g2.y};
    if (live) {
      fem::PlatePoint pts[2];
      lane_points(xl, half, pts[0], pts[1]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const fem::PlatePoint& g = pts[j];
        double wdx[4], wdy[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          wdx[a] = g.w * g.dx[a];
          wdy[a] = g.w * g.dy[a];
        }
        int s = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = a; b < 4; ++b, ++s) {
            I[s] = __builtin_fma(wdx[a], g.dx[b], I[s]);
            I[R_MYY + s] = __builtin_fma(wdy[a], g.dy[b], I[R_MYY + s]);
          }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) I[R_MXY + 4 * a + b] = __builtin_fma(wdx[a], g.dy[b], I[R_MXY + 4 * a + b]);
        const double al0 = g.al[0], al1 = g.al[2], be0 = g.be[0], be1 = g.be[1];
        const double wal0 = g.w * al0, wal1 = g.w * al1, wbe0 = g.w * be0, wbe1 = g.w * be1;
        I[R_A + 0] = __builtin_fma(wal0, al0, I[R_A + 0]);
        I[R_A + 1] = __builtin_fma(wal0, al1, I[R_A + 1]);
        I[R_A + 2] = __builtin_fma(wal1, al1, I[R_A + 2]);
        I[R_B + 0] = __builtin_fma(wbe0, be0, I[R_B + 0]);
        I[R_B + 1] = __builtin_fma(wbe0, be1, I[R_B + 1]);
        I[R_B + 2] = __builtin_fma(wbe1, be1, I[R_B + 2]);
      }
    }
    // points {0, 1} + points {2, 3}: lane half 0 writes integrals 0..20, half 1
    // integrals 21..41, so each lane sends the partner the partial the partner
    // writes and adds the one it receives to its own (IEEE addition commutes:
    // the sum is the same whichever lane forms it).  S[k] = integral k of this
    // lane's half (k + 21 h).
    double* r = rec + le * PL_REC;
    double S[PL_NINT / 2];
#pragma unroll
    for (int k = 0; k < PL_NINT / 2; ++k) {
      const double mine = half ? I[k + PL_NINT / 2] : I[k];
      const double theirs = swap_pair(half ? I[k] : I[k + PL_NINT / 2]);
      S[k] = mine + theirs;
    }
    if (live) {
      if (half == 0) {
#pragma unroll
        for (int k = 0; k < 20; k += 2) {
          *reinterpret_cast<double2*>(r + k) = make_double2(S[k], S[k + 1]);
        }
        r[20] = S[20];
        const fem::PlateMat mt = fem::plate_material(g3.x, g3.y, g4.x, g4.y);
        // bending constants are the membrane ones times t^2 / 12 (plate.rs:532-636)
        const double tb = (g4.x * g4.x) / 12.0;
        *reinterpret_cast<double2*>(r + R_MAT) = make_double2(mt.cm00, mt.cm01);
        *reinterpret_cast<double2*>(r + R_MAT + 2) = make_double2(mt.cm22, tb);
        *reinterpret_cast<double2*>(r + R_MAT + 4) = make_double2(mt.cs, 0.0);
      } else {
        r[21] = S[0];
#pragma unroll
        for (int k = 22; k < PL_NINT; k += 2) {
          *reinterpret_cast<double2*>(r + k) = make_double2(S[k - 21], S[k - 20]);
        }
        // shear edge vectors (plate.rs:452-508): e12, e43, e14, e23
        const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
        *reinterpret_cast<double2*>(r + R_SH) = make_double2(-(y1 - y2) * 0.25, (x1 - x2) * 0.25);
        *reinterpret_cast<double2*>(r + R_SH + 2) = make_double2(-(y4) * 0.25, (x4) * 0.25);
        *reinterpret_cast<double2*>(r + R_SH + 4) = make_double2(-(y1 - y4) * 0.25, (x1 - x4) * 0.25);
        *reinterpret_cast<double2*>(r + R_SH + 6) = make_double2(-(y2) * 0.25, (x2) * 0.25);
      }
    }
  }
  wave_lds_sync();

  // ---- step 2: trips over the lane's contributions.  The sum of the block
  // that item t belongs to, up to item t, is held in slot t (item t continues
  // the block of item t - 1 or starts a new one), so every slot index is a
  // compile-time constant and the finished blocks stay in registers until
  // every lane has read its last record.
  const uint64_t items = uint64_t(desc.x) | uint64_t(desc.y & 0xFFFFFu) << 32;
  const int cnt = int((desc.y >> 20) & 7u);
  double acc[PL_TRIPS][13];
  int run[PL_TRIPS];  // contributions summed into slot t
#pragma unroll
  for (int t = 0; t < PL_TRIPS; ++t) {
    run[t] = 0;
    if (t < trips && t < cnt) {
      const uint32_t it = uint32_t(items >> (13 * t)) & 0x1FFFu;
      const int e = int(it & 31u), p = int((it >> 5) & 3u), q = int((it >> 7) & 3u), k = int(it >> 9);
      const double* r = rec + e * PL_REC;
      const int si = sym_index(min(p, q), max(p, q));
      const double Mxx = r[si], Myy = r[R_MYY + si];
      const double Mxy = r[R_MXY + 4 * p + q], Myx = r[R_MXY + 4 * q + p];
      // alpha classes {0, 1} | {2, 3}; beta classes {0, 3} | {1, 2}
      const int bp = (p ^ (p >> 1)) & 1, bq = (q ^ (q >> 1)) & 1;
      const double Aa = r[R_A + (p >> 1) + (q >> 1)];
      const double Bb = r[R_B + bp + bq];
      // shear rows (plate.rs:452-508): v_a = (+-1/2, e), e = e12 for a in {0, 1},
      // e43 for {2, 3}, sign - for a in {1, 2}; u_a = (+-1/2, f), f = e14 for
      // a in {0, 3}, e23 for {1, 2}, sign - for a in {2, 3}
      const double2 evp = *reinterpret_cast<const double2*>(r + R_SH + 2 * (p >> 1));
      const double2 evq = *reinterpret_cast<const double2*>(r + R_SH + 2 * (q >> 1));
      const double2 eup = *reinterpret_cast<const double2*>(r + R_SH + 4 + 2 * bp);
      const double2 euq = *reinterpret_cast<const double2*>(r + R_SH + 4 + 2 * bq);
      const double2 m0 = *reinterpret_cast<const double2*>(r + R_MAT);
      const double2 m1 = *reinterpret_cast<const double2*>(r + R_MAT + 2);
      const double cs = r[R_MAT + 4];
      double o[13];
      block_entries(Mxx, Mxy, Myx, Myy, Aa * cs, Bb * cs, m0.x, m0.y, m1.x, m1.y, evp, eup, evq, euq, bp, p >> 1,
                    bq, q >> 1, o);
      if (t == 0) {
        run[0] = 1;
#pragma unroll
        for (int i = 0; i < 13; ++i) acc[0][i] = o[i];
      } else {
        const bool first = item_block(items, t - 1) != k;
        run[t] = first ? 1 : run[t - 1] + 1;
        const double keep = first ? 0.0 : 1.0;  // acc * 1 + o is acc + o exactly
#pragma unroll
        for (int i = 0; i < 13; ++i) acc[t][i] = __builtin_fma(acc[t - 1][i], keep, o[i]);
      }
    }
  }
  wave_lds_sync();  // every record read is done: the staging reuses the records' LDS

  // finished blocks (the last item of each block) into the node's staging segment
  {
    const int slot = int((desc.y >> 23) & 15u);
    const int nb = int((desc.y >> 27) & 15u);
    // staging position of the lane's node, from the node table lane `slot`
    const int pos = __builtin_amdgcn_ds_bpermute(slot << 2, int(nd.y >> 18));
    double* seg = stage + pos;
#pragma unroll
    for (int t = 0; t < PL_TRIPS; ++t) {
      if (t < trips && t < cnt && (t + 1 == cnt || item_block(items, t + 1) != item_block(items, t))) {
        const uint32_t it = uint32_t(items >> (13 * t)) & 0x1FFFu;
        const int k = int(it >> 9);
        double* x = seg + 2 * k;
        x[0] = acc[t][0];
        x[1] = acc[t][1];
        x[2 * nb] = acc[t][2];
        x[2 * nb + 1] = acc[t][3];
        double* z = seg + 4 * nb + 3 * k;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
          z[3 * nb * rr] = acc[t][4 + 3 * rr];
          z[3 * nb * rr + 1] = acc[t][5 + 3 * rr];
          z[3 * nb * rr + 2] = acc[t][6 + 3 * rr];
        }
        if (((it >> 5) & 3u) == ((it >> 7) & 3u)) {
          // drilling diagonal: KROT6 of every element of the node, in insertion order
          double dr = fem::KROT6;
          for (int m = 1; m < run[t]; ++m) dr += fem::KROT6;
          seg[13 * nb] = dr;
        }
      }
    }
  }
  wave_lds_sync();

  // ---- step 3: copy-out by runs of owned nodes whose CSR ranges follow each
  // other (staged contiguously, at the CSR parity): consecutive lanes store
  // consecutive aligned 16-byte chunks, the at most two half chunks at a run's
  // ends as 8-byte stores; non-temporal (nothing in this kernel reads the
  // values again)
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A.val, 0, int(A.val_bytes), 0x00020000);
  int s0 = 0;
  while (s0 < nnod) {
    const uint32_t base = uint32_t(__builtin_amdgcn_readlane(int(nd.x), s0));
    const uint32_t sp = uint32_t(__builtin_amdgcn_readlane(int(nd.y), s0)) >> 18;  // staging position of base
    uint32_t end = base + 13u * (uint32_t(__builtin_amdgcn_readlane(int(nd.y), s0)) & 15u) + 1u;
    int s1 = s0 + 1;
    while (s1 < nnod && uint32_t(__builtin_amdgcn_readlane(int(nd.x), s1)) == end) {
      end += 13u * (uint32_t(__builtin_amdgcn_readlane(int(nd.y), s1)) & 15u) + 1u;
      ++s1;
    }
    const int shift = int(sp >> 1) - int(base >> 1);  // staging chunk - CSR chunk (same parity)
    // lane l of a store covers chunk l of a 128-byte-aligned group of 64 chunks
    for (uint32_t c = ((base >> 1) & ~7u) + uint32_t(l); c < (end + 1u) >> 1; c += 64u) {
      if (c < (base >> 1)) continue;
      const double2 v = stage2[int(c) + shift];
      const uint32_t g = 2u * c;  // first value of the chunk
      if (g >= base && g + 1u < end) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(pl_u32x4, v), rsrc, 8u * g, 0, PL_STORE_NT);
      } else if (g < base) {  // the run starts at an odd position: the chunk's upper half
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pl_u32x2, v.y), rsrc, 8u * g + 8u, 0, PL_STORE_NT);
      } else {  // the run ends at an odd position: the chunk's lower half
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pl_u32x2, v.x), rsrc, 8u * g, 0, PL_STORE_NT);
      }
    }
    s0 = s1;
  }
}

std::string plate_lanes_gather(Device* d, hipStream_t s) {
  const size_t n_slots = size_t(d->pl_waves) * PL_ELEMS;
  if (n_slots) {
    plate_lanes_gather_kernel<<<unsigned((n_slots + 255) / 256), 256, 0, s>>>(d->xyz, d->params[PLATE], d->pb_conn,
                                                                            d->pb_eid, n_slots,
                                                                            d->pl_uniform ? PL_GEO_U : PL_GEO, d->pl_geo);
    HIP_TRY(hipGetLastError());
  }
  d->pl_geo_dirty = false;
  return "";
}

}  // namespace

std::string upload_plate_lane_maps(Device* d, const PlateLaneMaps& pm) {
  d->pl_waves = pm.n_waves;
  std::string e;
  if (!(e = dupload(d->pb_conn, pm.sconn.data(), pm.sconn.size())).empty()) return e;
  if (!(e = dupload(d->pb_eid, pm.seid.data(), pm.seid.size())).empty()) return e;
  if (!(e = dupload(d->pl_nodes, pm.nodes.data(), pm.nodes.size())).empty()) return e;
  if (!(e = dupload(d->pl_lanes, pm.lanes.data(), pm.lanes.size())).empty()) return e;
  if (!(e = dupload(d->pl_lane_set, pm.lane_set.data(), pm.lane_set.size())).empty()) return e;
  if (!(e = dalloc(d->pl_geo, size_t(pm.n_waves) * PL_ELEMS * PL_GEO)).empty()) return e;  // either layout fits
  e = plate_lanes_gather(d, nullptr);
  if (!e.empty()) return e;
  HIP_TRY(hipDeviceSynchronize());
  return "";
}

std::string set_plate_lane_material(Device* d, bool uniform, const double mat[3]) {
  d->pl_uniform = uniform;
  for (int k = 0; k < 3; ++k) d->pl_mat[k] = uniform ? mat[k] : 0.0;
  d->pl_geo_dirty = true;
  return "";
}

std::string assemble_plate_lanes(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (d->pl_geo_dirty) {  // coordinates or parameters changed since the slots were filled
    auto e = plate_lanes_gather(d, s);
    if (!e.empty()) return e;
  }
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->pl_waves > 0) {
    LaneArgs a;
    a.geo = d->pl_geo;
    a.nodes = reinterpret_cast<const uint2*>(d->pl_nodes);
    a.lanes = reinterpret_cast<const uint2*>(d->pl_lanes);
    a.lane_set = d->pl_lane_set;
    a.geo_stride = uint32_t(d->pl_uniform ? PL_GEO_U : PL_GEO);
    for (int k = 0; k < 3; ++k) a.mat[k] = d->pl_mat[k];
    a.val = val;
    a.val_bytes = uint32_t(d->nnz * 8);
    plate_lanes_kernel<<<d->pl_waves, 64, 0, s>>>(a);
    HIP_TRY(hipGetLastError());
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

}  // namespace femdev
