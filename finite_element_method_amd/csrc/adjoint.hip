// adjoint.hip -- backward (adjoint) kernels of the assembly: derivatives of a
// scalar L(K) with respect to element parameters and node coordinates, given
// dL/dK on the CSR pattern, and the compliance sensitivities -u_e^T (dK_e/dp) u_e
// given nodal displacements u.  Element stiffness is recomputed, not stored.
//
// Element gradient blocks are gathered through the pattern's block structure:
// element node pair (p, q) -> global node-pair block k (bid), whose CSR entries
// in row i start at row_ptr[6 a + i] + boff[6 k + i] and follow the block's
// union mask bmask[k].  All pattern entries of the block are read (not only the
// element's own structural entries), so coordinate derivatives that create
// entries the element lacks at its current geometry are still seen.
//
// Determinism: per-element results are written once; the four lanes of a plate
// element combine with a fixed butterfly; nodal coordinate gradients are
// per-(element, node) rows, written to node-major slots and summed per node in
// element-insertion order by node_gather_kernel.  No atomics.
#include <hip/hip_runtime.h>

#include <string>

#include "bar_adjoint.hpp"
#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"
#include "lane_ops.hpp"
#include "plate_adjoint.hpp"

namespace femdev {

namespace {

constexpr int MODE_DLDK = 0;  // g = dL/dK values (nnz)
constexpr int MODE_U = 1;     // g = nodal displacements (6 N); L = -u^T K u per element

struct AdjArgs {
  const double* xyz;
  const uint32_t* conn;
  const double* params;
  const double* R;
  const double* derived;
  size_t n;
  const double* g;
  const uint32_t* bid;
  const uint64_t* bmask;
  const uint16_t* boff;
  // flat plates, per (element, p): CSR position of row 6 a of corner p's node a,
  // and the block index k of each q (6 bits each) | nb << 24
  const uint2* kpack;
  const uint32_t* eorder;  // flat plates: the elements in workgroup order
  const int64_t* row_ptr;
  double* dparam;  // n x params_per, may be null
  double* xs;      // coordinate-gradient rows in node-major slots (3 doubles each), may be null
  const uint32_t* xpos;  // per (element, node) of this type: its slot in xs
  double rel_tol;
  uint32_t g_bytes;  // dL/dK: bytes of g when below 4 GiB (wide range-checked loads), else 0
};

// logical block of hardware block b: the hardware deals blocks to the eight
// XCDs round robin (b % 8); each XCD gets a contiguous range of element blocks
// so that neighbouring elements share an L2
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned x = b & 7u, j = b >> 3, q = nb >> 3, r = nb & 7u;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// global-frame 6x6 gradient block of node pair block k, rows of node a.  The
// block's entries of row i are contiguous in the CSR row: load the run, then
// place it by the row mask.
// With g_bytes set the run is read in 16-byte pairs (and one 8-byte tail) by
// buffer loads, range-checked against g: half the memory requests of one load
// per value, and the requests, not the bytes, bound this gather.
__device__ inline void gather_block(const AdjArgs& A, uint32_t k, uint32_t a, double G[36]) {
  const uint64_t um = A.bmask[k];
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(A.g), 0, int(A.g_bytes), 0x00020000);
  uint32_t bo[3];  // the block's six row offsets, two per word (boff rows start 4-byte aligned)
#pragma unroll
  for (int m = 0; m < 3; ++m) bo[m] = reinterpret_cast<const uint32_t*>(A.boff + 6 * size_t(k))[m];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const uint32_t mi = uint32_t(um >> (6 * i)) & 63u;
    const int n = __builtin_popcount(mi);
    const uint32_t off_i = (bo[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
    const int64_t pos = A.row_ptr[6 * size_t(a) + i] + off_i;
    double v[6];
    if (A.g_bytes) {
      const uint32_t byte = uint32_t(pos) * 8u;
#pragma unroll
      for (int c = 0; c < 6; c += 2) {
        v[c] = 0.0;
        v[c + 1] = 0.0;
        if (c + 1 < n) {
          const double2 w = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte + 8u * c, 0, 0));
          v[c] = w.x;
          v[c + 1] = w.y;
        } else if (c < n) {
          v[c] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, byte + 8u * c, 0, 0));
        }
      }
    } else {
      const double* src = A.g + pos;
#pragma unroll
      for (int c = 0; c < 6; ++c) v[c] = c < n ? src[c] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int c = __builtin_popcount(mi & ((1u << j) - 1u));
      double x = v[0];
#pragma unroll
      for (int cc = 1; cc <= j; ++cc) x = c == cc ? v[cc] : x;
      G[i * 6 + j] = ((mi >> j) & 1u) ? x : 0.0;
    }
  }
}

__device__ inline bool is_identity(const double R[9]) {
  return R[0] == 1.0 && R[4] == 1.0 && R[8] == 1.0 && R[1] == 0.0 && R[2] == 0.0 && R[3] == 0.0 && R[5] == 0.0 &&
         R[6] == 0.0 && R[7] == 0.0;
}

// Gl = R6 G R6^T in place (sub-blocks R G_XY R^T)
__device__ inline void to_local(double G[36], const double R[9]) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      double T[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          T[i * 3 + j] = (R[i * 3 + 0] * G[(3 * bi + 0) * 6 + 3 * bj + j] + R[i * 3 + 1] * G[(3 * bi + 1) * 6 + 3 * bj + j]) +
                         R[i * 3 + 2] * G[(3 * bi + 2) * 6 + 3 * bj + j];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          G[(3 * bi + i) * 6 + 3 * bj + j] = (T[i * 3 + 0] * R[j * 3 + 0] + T[i * 3 + 1] * R[j * 3 + 1]) + T[i * 3 + 2] * R[j * 3 + 2];
    }
}

// local-frame displacement of node a: (R u_t, R u_r)
__device__ inline void local_u(const double* __restrict__ u, uint32_t a, const double R[9], double o[6]) {
  double v[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = u[6 * size_t(a) + i];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      o[3 * h + i] = (R[i * 3 + 0] * v[3 * h + 0] + R[i * 3 + 1] * v[3 * h + 1]) + R[i * 3 + 2] * v[3 * h + 2];
}

template <int CTRL>
__device__ inline double quad_perm(double v) {
  typedef unsigned int qp_u32x2 __attribute__((ext_vector_type(2)));
  const qp_u32x2 x = __builtin_bit_cast(qp_u32x2, v);
  qp_u32x2 r;
  r.x = uint32_t(__builtin_amdgcn_mov_dpp(int(x.x), CTRL, 0xF, 0xF, false));
  r.y = uint32_t(__builtin_amdgcn_mov_dpp(int(x.y), CTRL, 0xF, 0xF, false));
  return __builtin_bit_cast(double, r);
}
__device__ inline double quad_sum(double v) {
  v += quad_perm<0xB1>(v);
  v += quad_perm<0x4E>(v);
  return v;
}

// ---------------------------------------------------------------- plates
// point data of the block's 64 elements: lane p of an element computes point p
constexpr int PA_POINT = 9;  // double2 per (element, point): plate_point's 17 values and a pad
__device__ inline void store_point(double2* dst, const fem::PlatePoint& g) {
  const double v[18] = {g.w, g.dx[0], g.dx[1], g.dx[2], g.dx[3], g.dy[0], g.dy[1], g.dy[2], g.dy[3],
                        g.al[0], g.al[1], g.al[2], g.al[3], g.be[0], g.be[1], g.be[2], g.be[3], 0.0};
#pragma unroll
  for (int k = 0; k < PA_POINT; ++k) dst[k] = make_double2(v[2 * k], v[2 * k + 1]);
}
__device__ inline fem::PlatePoint load_point(const double2* src) {
  double v[18];
#pragma unroll
  for (int k = 0; k < PA_POINT; ++k) {
    const double2 x = src[k];
    v[2 * k] = x.x;
    v[2 * k + 1] = x.y;
  }
  fem::PlatePoint g;
  g.w = v[0];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    g.dx[a] = v[1 + a];
    g.dy[a] = v[5 + a];
    g.al[a] = v[9 + a];
    g.be[a] = v[13 + a];
  }
  return g;
}

// element data of plate e: nodes, frame, local coordinates relative to node 3
// (xl = x, y of nodes 1, 2, 4), parameters
struct PlateElem {
  uint32_t nd[4];
  double R[9];
  double loc[4][3];
  double xl[6];
  double E, nu, t, k;
};

// FLAT: every plate frame is the identity (constants instead of loads)
template <bool FLAT>
__device__ inline void load_plate_elem(const AdjArgs& A, size_t e, PlateElem& o) {
#pragma unroll
  for (int a = 0; a < 4; ++a) o.nd[a] = A.conn[4 * e + a];
  if (FLAT) {
#pragma unroll
    for (int k = 0; k < 9; ++k) o.R[k] = (k % 4 == 0) ? 1.0 : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) o.R[k] = A.R[9 * e + k];
  }
  const size_t n3 = 3 * size_t(o.nd[2]);
  const double x3 = A.xyz[n3], y3 = A.xyz[n3 + 1], z3 = A.xyz[n3 + 2];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const size_t na = 3 * size_t(o.nd[a]);
    const double dd[3] = {A.xyz[na] - x3, A.xyz[na + 1] - y3, A.xyz[na + 2] - z3};
#pragma unroll
    for (int i = 0; i < 3; ++i)
      o.loc[a][i] = (o.R[i * 3 + 0] * dd[0] + o.R[i * 3 + 1] * dd[1]) + o.R[i * 3 + 2] * dd[2];
    if (FLAT) o.loc[a][2] = 0.0;
  }
  o.xl[0] = o.loc[0][0]; o.xl[1] = o.loc[0][1];
  o.xl[2] = o.loc[1][0]; o.xl[3] = o.loc[1][1];
  o.xl[4] = o.loc[3][0]; o.xl[5] = o.loc[3][1];
  o.E = A.params[4 * e];
  o.nu = A.params[4 * e + 1];
  o.t = A.params[4 * e + 2];
  o.k = A.params[4 * e + 3];
}

// sums over the four lanes of an element (dead lanes take part), then lane p
// writes parameter p and node p's coordinate gradient
__device__ inline void plate_write(const AdjArgs& A, size_t e, int p, bool live, const PlateElem& el, double P[7],
                                   double gxl[6], double hx, double hy) {
#pragma unroll
  for (int i = 0; i < 7; ++i) P[i] = quad_sum(P[i]);
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = quad_sum(gxl[i]);
  hx = quad_sum(hx);
  hy = quad_sum(hy);
  if (!live) return;
  double gp[4], gX[4][3];
  fem::plate_adjoint_finish(P, gxl, hx, hy, el.loc, el.R, el.E, el.nu, el.t, el.k, gp, gX);
  if (A.dparam) A.dparam[4 * e + p] = fem::sel4(gp[0], gp[1], gp[2], gp[3], p);
  if (A.xs) {
    double* x = A.xs + 3 * size_t(A.xpos[4 * e + p]);
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = fem::sel4(gX[0][i], gX[1][i], gX[2][i], gX[3][i], p);
  }
}

// Plates in any frame: four lanes per element (lane p = row block of node p),
// 64 elements per block; gradient blocks gathered through the block maps.
__global__ __launch_bounds__(256) void plate_adjoint_kernel(AdjArgs A) {
  __shared__ double2 pts[64 * 4 * PA_POINT];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t e0 = size_t(b) * 64 + (threadIdx.x >> 2);
  const int p = int(threadIdx.x & 3u);
  const bool live = e0 < A.n;
  const size_t e = live ? e0 : A.n - 1;  // dead lanes still take part in the lane sums
  PlateElem el;
  load_plate_elem<false>(A, e, el);
  const uint32_t* nd = el.nd;
  const double* R = el.R;
  const bool ident = is_identity(R);
  const double* xl = el.xl;
  const fem::PlateMat mat = fem::plate_material(el.E, el.nu, el.t, el.k);
  fem::PlateShear sv;
  fem::plate_shear_vectors(xl, sv);
  const uint32_t ap = p == 0 ? nd[0] : (p == 1 ? nd[1] : (p == 2 ? nd[2] : nd[3]));
  // pass 1: integrals of row p; the point data of point p come from lane p
  double2* my_pts = pts + (threadIdx.x >> 2) * 4 * PA_POINT;
  {
    fem::PlatePoint g;
    fem::plate_point(xl, p, g);
    store_point(my_pts + p * PA_POINT, g);
  }
  __syncthreads();
  double acc[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) acc[i] = 0.0;
#pragma unroll 1
  for (int ip = 0; ip < 4; ++ip) {
    const fem::PlatePoint g = load_point(my_pts + ip * PA_POINT);
    const double wdx = g.w * fem::sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p);
    const double wdy = g.w * fem::sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
    const double wal = g.w * fem::sel4(g.al[0], g.al[1], g.al[2], g.al[3], p);
    const double wbe = g.w * fem::sel4(g.be[0], g.be[1], g.be[2], g.be[3], p);
    fem::plate_row_accumulate(g, wdx, wdy, wal, wbe, acc);
  }
  // pass 2: gradient blocks of row p.  acc[] doubles as c[]: column q's
  // integral coefficients replace its integrals once they are consumed.
  double* c = acc;
  double P[7], gxl[6], hx = 0.0, hy = 0.0;
#pragma unroll
  for (int i = 0; i < 7; ++i) P[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = 0.0;
  double vp_[3], up_[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vp_[i] = fem::sel4(sv.v[0][i], sv.v[1][i], sv.v[2][i], sv.v[3][i], p);
    up_[i] = fem::sel4(sv.u[0][i], sv.u[1][i], sv.u[2][i], sv.u[3][i], p);
  }
  // one block per iteration (q is wave-uniform): only one block's temporaries live
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    double gl[36];
    gather_block(A, A.bid[16 * e + 4 * p + q], ap, gl);
    if (!ident) to_local(gl, R);
    double Mq[6], vq[3], uq_[3], cq[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Mq[k] = fem::sel4(acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3], q);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      vq[i] = fem::sel4(sv.v[0][i], sv.v[1][i], sv.v[2][i], sv.v[3][i], q);
      uq_[i] = fem::sel4(sv.u[0][i], sv.u[1][i], sv.u[2][i], sv.u[3][i], q);
    }
    fem::plate_block_adjoint_q(Mq, mat, vp_, up_, vq, uq_, p, q, gl, P, cq, gxl, hx, hy, true);
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) c[4 * k + j] = j == q ? cq[k] : c[4 * k + j];
  }
  // pass 3: reverse sweep over the integration points.  The sweep is linear in
  // the seven adjoints of a point's forward quantities: the four rows' values
  // are added first, and lane p sweeps point p.
  {
    double mine[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int ip = 0; ip < 4; ++ip) {
      double s7[7];
      fem::plate_point_scalars(load_point(my_pts + ip * PA_POINT), ip, p, c, s7);
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const double tot = quad_sum(s7[k]);
        mine[k] = ip == p ? tot : mine[k];
      }
    }
    fem::plate_point_reverse(xl, p, mine, gxl);
  }
  plate_write(A, e, p, live, el, P, gxl, hx, hy);
}

// plate_write for a flat plate (identity frame, nodes in z = 0, no frame term)
__device__ inline void plate_write_flat(const AdjArgs& A, size_t e, int p, bool live, double E, double nu, double t,
                                        double k, double P[7], double gxl[6]) {
#pragma unroll
  for (int i = 0; i < 7; ++i) P[i] = quad_sum(P[i]);
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = quad_sum(gxl[i]);
  if (!live) return;
  double gp[4], gX[4][3];
  fem::plate_adjoint_finish_flat(P, gxl, E, nu, t, k, gp, gX);
  if (A.dparam) A.dparam[4 * e + p] = fem::sel4(gp[0], gp[1], gp[2], gp[3], p);
  if (A.xs) {
    double* x = A.xs + 3 * size_t(A.xpos[4 * e + p]);
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = fem::sel4(gX[0][i], gX[1][i], gX[2][i], gX[3][i], p);
  }
}

// ------------------------------------------------ flat plates, dL/dK
// One wave per block, 16 elements, four lanes per element.  Lane p is row
// block p while the gradient blocks are read and integration point p while
// the integrals are swept back; the element's data meet in LDS:
//   pts:  the four points' data (w = det J, dh/dx, dh/dy, the two distinct
//         MITC4 multipliers of each direction);
//   rows: per row p and column q six values, first the integrals (Mxx, Mxy,
//         Myx, Myy, A, B), then their coefficients dL/d(integral).
// The sweep of point ip needs only the sums of the (p, q) and (q, p)
// coefficients, and the point's value adjoint is half the gradient vectors'
// product with the point data (the integrand is a quadratic form in them).
constexpr int PF_PT = 14;              // doubles per point
constexpr int PF_PTS = 4 * PF_PT + 2;  // per element, padded so that elements start in different banks
constexpr int PF_ROWS = 4 * 24 + 2;    // per element: 4 rows x 4 columns x 6 values, padded

// compiler fence: the accumulators are final at it and LDS reads after it
// are not issued before it, so the coefficients in registers stay few
__device__ inline void sweep_fence(double g[16]) {
#pragma unroll
  for (int i = 0; i < 16; i += 4) asm volatile("" : "+v"(g[i]), "+v"(g[i + 1]), "+v"(g[i + 2]), "+v"(g[i + 3])::"memory");
}

// block k of a flat plate node whose rows' values start at gr (nb blocks per
// row): 2 values per row in the two membrane rows (at 2 k and 2 nb + 2 k), 3
// per row in the three bending rows (at (4 + 3 i) nb + 3 k), then, on the
// diagonal block only, the drilling entry (at 13 nb).  The drilling entry has
// no derivative: 13 values are loaded, (u,u) (u,v) (v,u) (v,v) and the
// (w, theta_x, theta_y) x (w, theta_x, theta_y) block row-major.
__device__ inline void load_flat_k13(const double* gr, int nb, int k, double v[13]) {
  v[0] = gr[2 * k];
  v[1] = gr[2 * k + 1];
  v[2] = gr[2 * nb + 2 * k];
  v[3] = gr[2 * nb + 2 * k + 1];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) v[4 + 3 * i + j] = gr[(4 + 3 * i) * nb + 3 * k + j];
}

// shear-vector adjoint of node a given as masks (a varies across lanes): v_a
// carries e12 for a <= 1, e43 otherwise; u_a carries e14 for a = 0, 3, e23 otherwise
__device__ inline void shear_vector_adjoint_lane(int a, double gv1, double gv2, double gu1, double gu2,
                                                 double gxl[6]) {
  const double q = 0.25;
  const double v12 = a <= 1 ? q : 0.0, v43 = a <= 1 ? 0.0 : q;
  const double u14 = (a == 0 || a == 3) ? q : 0.0, u23 = (a == 0 || a == 3) ? 0.0 : q;
  gxl[0] += gv2 * v12 + gu2 * u14;
  gxl[1] -= gv1 * v12 + gu1 * u14;
  gxl[2] += gu2 * u23 - gv2 * v12;
  gxl[3] += gv1 * v12 - gu1 * u23;
  gxl[4] += gv2 * v43 - gu2 * u14;
  gxl[5] += gu1 * u14 - gv1 * v43;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void plate_adjoint_flat_kernel(AdjArgs A) {
  __shared__ double2 pts[16 * PF_PTS / 2];
  __shared__ double2 rows[16 * PF_ROWS / 2];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const int le = int(threadIdx.x >> 2);
  const int p = int(threadIdx.x & 3u);
  const size_t i0 = size_t(b) * 16 + le;
  const bool live = i0 < A.n;
  const size_t e = A.eorder[live ? i0 : A.n - 1];  // dead lanes still take part in the lane sums
  // gradient-block positions first: the value loads wait on this one load only
  const uint2 kr = A.kpack[4 * e + p];
  const uint32_t g0 = kr.x, kp = kr.y;
  const int nb = int(kp >> 24);
  double nxt[13], nxt2[13];
  load_flat_k13(A.g + g0, nb, int(kp & 63u), nxt);
  load_flat_k13(A.g + g0, nb, int((kp >> 6) & 63u), nxt2);
  double xl[6];
  fem::PlateMat mat;
  {
    PlateElem el;
    load_plate_elem<true>(A, e, el);
#pragma unroll
    for (int i = 0; i < 6; ++i) xl[i] = el.xl[i];
    mat = fem::plate_material(el.E, el.nu, el.t, el.k);
  }
  double* const ep = reinterpret_cast<double*>(pts) + le * PF_PTS;
  double* const er = reinterpret_cast<double*>(rows) + le * PF_ROWS;
  // point p's data; the reciprocals its reverse sweep needs stay with lane p
  double rdet, rqr, rqs;
  {
    fem::PlatePoint g;
    fem::plate_point_x(xl, p, g, rdet, rqr, rqs);
    double2* d = reinterpret_cast<double2*>(ep + p * PF_PT);
    d[0] = make_double2(g.w, g.dx[0]);
    d[1] = make_double2(g.dx[1], g.dx[2]);
    d[2] = make_double2(g.dx[3], g.dy[0]);
    d[3] = make_double2(g.dy[1], g.dy[2]);
    d[4] = make_double2(g.dy[3], g.al[0]);
    d[5] = make_double2(g.al[2], g.be[0]);
    d[6] = make_double2(g.be[1], 0.0);
  }
  __syncthreads();
  // integrals of row p (plate_row_accumulate with the multipliers' classes:
  // al_0 = al_1, al_2 = al_3, be_0 = be_3, be_1 = be_2)
  {
    double acc[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) acc[i] = 0.0;
#pragma unroll
    for (int ip = 0; ip < 4; ++ip) {
      const double2* d = reinterpret_cast<const double2*>(ep + ip * PF_PT);
      const double2 d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4], d5 = d[5], d6 = d[6];
      const double w = d0.x;
      const double dx[4] = {d0.y, d1.x, d1.y, d2.x}, dy[4] = {d2.y, d3.x, d3.y, d4.x};
      const double al[4] = {d4.y, d4.y, d5.x, d5.x}, be[4] = {d5.y, d6.x, d6.x, d5.y};
      const double wdx = w * ep[ip * PF_PT + 1 + p], wdy = w * ep[ip * PF_PT + 5 + p];
      const double wal = w * (p <= 1 ? al[0] : al[2]), wbe = w * ((p == 0 || p == 3) ? be[0] : be[1]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[q] = __builtin_fma(wdx, dx[q], acc[q]);
        acc[4 + q] = __builtin_fma(wdx, dy[q], acc[4 + q]);
        acc[8 + q] = __builtin_fma(wdy, dx[q], acc[8 + q]);
        acc[12 + q] = __builtin_fma(wdy, dy[q], acc[12 + q]);
        acc[16 + q] = __builtin_fma(wal, al[q], acc[16 + q]);
        acc[20 + q] = __builtin_fma(wbe, be[q], acc[20 + q]);
      }
    }
    double2* r = reinterpret_cast<double2*>(er + p * 24);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      r[3 * q] = make_double2(acc[q], acc[4 + q]);
      r[3 * q + 1] = make_double2(acc[8 + q], acc[12 + q]);
      r[3 * q + 2] = make_double2(acc[16 + q], acc[20 + q]);
    }
  }
  // the element's shear-vector edges (plate_shear_vectors)
  const double e12x = -(xl[1] - xl[3]) * 0.25, e12y = (xl[0] - xl[2]) * 0.25;
  const double e43x = -(xl[5]) * 0.25, e43y = (xl[4]) * 0.25;
  const double e14x = -(xl[1] - xl[5]) * 0.25, e14y = (xl[0] - xl[4]) * 0.25;
  const double e23x = -(xl[3]) * 0.25, e23y = (xl[2]) * 0.25;
  const double vp[3] = {(p == 0 || p == 3) ? 0.5 : -0.5, p <= 1 ? e12x : e43x, p <= 1 ? e12y : e43y};
  const double up[3] = {p <= 1 ? 0.5 : -0.5, (p == 0 || p == 3) ? e14x : e23x, (p == 0 || p == 3) ? e14y : e23y};
  double P[7], gxl[6];
#pragma unroll
  for (int i = 0; i < 7; ++i) P[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = 0.0;
  double gv1 = 0.0, gv2 = 0.0, gu1 = 0.0, gu2 = 0.0;  // dL/d(v_p), dL/d(u_p), components 1 and 2
  // blocks (p, q): q is wave-uniform; block q + 1's values load during block q
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    double G[13];
#pragma unroll
    for (int i = 0; i < 13; ++i) {
      G[i] = nxt[i];
      nxt[i] = nxt2[i];
    }
    if (q < 2) load_flat_k13(A.g + g0, nb, int((kp >> (6 * (q + 2))) & 63u), nxt2);
    double2* r = reinterpret_cast<double2*>(er + p * 24 + 6 * q);
    const double2 m0 = r[0], m1 = r[1], m2 = r[2];
    const double Mxx = m0.x, Mxy = m0.y, Myx = m1.x, Myy = m1.y, Aq = m2.x, Bq = m2.y;
    const double vq[3] = {(q == 0 || q == 3) ? 0.5 : -0.5, q <= 1 ? e12x : e43x, q <= 1 ? e12y : e43y};
    const double uq[3] = {q <= 1 ? 0.5 : -0.5, (q == 0 || q == 3) ? e14x : e23x, (q == 0 || q == 3) ? e14y : e23y};
    const double Guu = G[0], Guv = G[1], Gvu = G[2], Gvv = G[3];
    const double Gxx = G[8], Gxy = G[9], Gyx = G[11], Gyy = G[12];
    // constitutive partials
    P[0] += Guu * Mxx + Gvv * Myy;
    P[1] += Guv * Mxy + Gvu * Myx;
    P[2] += ((Guu * Myy + Guv * Myx) + Gvu * Mxy) + Gvv * Mxx;
    P[3] += Gxx * Myy + Gyy * Mxx;
    P[4] -= Gxy * Myx + Gyx * Mxy;
    P[5] += ((Gxx * Mxx - Gxy * Mxy) - Gyx * Myx) + Gyy * Myy;
    // shear block Gs = G[4..12] over (w, theta_x, theta_y)
    double Gv[3], Gu[3], vG[3], uG[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      Gv[i] = (G[4 + 3 * i] * vq[0] + G[5 + 3 * i] * vq[1]) + G[6 + 3 * i] * vq[2];
      Gu[i] = (G[4 + 3 * i] * uq[0] + G[5 + 3 * i] * uq[1]) + G[6 + 3 * i] * uq[2];
      vG[i] = (vp[0] * G[4 + i] + vp[1] * G[7 + i]) + vp[2] * G[10 + i];
      uG[i] = (up[0] * G[4 + i] + up[1] * G[7 + i]) + up[2] * G[10 + i];
    }
    const double sv_ = (vp[0] * Gv[0] + vp[1] * Gv[1]) + vp[2] * Gv[2];
    const double su_ = (up[0] * Gu[0] + up[1] * Gu[1]) + up[2] * Gu[2];
    P[6] += Aq * sv_ + Bq * su_;
    // coefficients of column q's integrals
    const double cxx = ((mat.cm00 * Guu + mat.cm22 * Gvv) + mat.cb22 * Gxx) + mat.cb00 * Gyy;
    const double cxy = ((mat.cm01 * Guv + mat.cm22 * Gvu) - mat.cb22 * Gxy) - mat.cb01 * Gyx;
    const double cyx = ((mat.cm22 * Guv + mat.cm01 * Gvu) - mat.cb01 * Gxy) - mat.cb22 * Gyx;
    const double cyy = ((mat.cm22 * Guu + mat.cm00 * Gvv) + mat.cb00 * Gxx) + mat.cb22 * Gyy;
    r[0] = make_double2(cxx, cxy);
    r[1] = make_double2(cyx, cyy);
    r[2] = make_double2(mat.cs * sv_, mat.cs * su_);
    // shear vectors: d/dv_p = cs A Gs v_q, d/dv_q = cs A v_p^T Gs (same with u, B)
    const double sa = mat.cs * Aq, sb = mat.cs * Bq;
    gv1 += sa * Gv[1];
    gv2 += sa * Gv[2];
    gu1 += sb * Gu[1];
    gu2 += sb * Gu[2];
    shear_vector_adjoint_lane(q, sa * vG[1], sa * vG[2], sb * uG[1], sb * uG[2], gxl);
  }
  shear_vector_adjoint_lane(p, gv1, gv2, gu1, gu2, gxl);
  __syncthreads();
  // sweep of point p over the sums of the (a, b) and (b, a) coefficients:
  // gdx_a = dl/d(dx_a), ... of l = w * sum_ab [dx_a cxx dx_b + dx_a cxy dy_b
  // + dy_a cyx dx_b + dy_a cyy dy_b + al_a cA al_b + be_a cB be_b]
  {
    const double2* d = reinterpret_cast<const double2*>(ep + p * PF_PT);
    const double2 d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4], d5 = d[5], d6 = d[6];
    const double w = d0.x;
    const double dx[4] = {d0.y, d1.x, d1.y, d2.x}, dy[4] = {d2.y, d3.x, d3.y, d4.x};
    const double al[4] = {d4.y, d4.y, d5.x, d5.x}, be[4] = {d5.y, d6.x, d6.x, d5.y};
    double gv[16];  // dl/d(dx_a), dl/d(dy_a), dl/d(al_a), dl/d(be_a)
    double* const gdx = gv;
    double* const gdy = gv + 4;
    double* const gal = gv + 8;
    double* const gbe = gv + 12;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double2* c = reinterpret_cast<const double2*>(er + a * 24 + 6 * a);
      const double2 c0 = c[0], c1 = c[1], c2 = c[2];
      const double sxy = c0.y + c1.x;
      gdx[a] = 2.0 * c0.x * dx[a] + sxy * dy[a];
      gdy[a] = sxy * dx[a] + 2.0 * c1.y * dy[a];
      gal[a] = 2.0 * c2.x * al[a];
      gbe[a] = 2.0 * c2.y * be[a];
    }
    // pairs in three groups of two, each group's reads issued after the
    // previous group's use (all 96 coefficients at once would not fit)
#pragma unroll
    for (int pr = 0; pr < 6; ++pr) {
      const int a = pr < 3 ? 0 : (pr < 5 ? 1 : 2);
      const int bb = pr < 3 ? pr + 1 : (pr < 5 ? pr - 1 : 3);
      if ((pr & 1) == 0) sweep_fence(gv);
      {
        const double2* cab = reinterpret_cast<const double2*>(er + a * 24 + 6 * bb);
        const double2* cba = reinterpret_cast<const double2*>(er + bb * 24 + 6 * a);
        const double2 x0 = cab[0], x1 = cab[1], x2 = cab[2], y0 = cba[0], y1 = cba[1], y2 = cba[2];
        const double sxx = x0.x + y0.x;   // dx_a dx_b
        const double sab = x0.y + y1.x;   // dx_a dy_b
        const double sba = x1.x + y0.y;   // dy_a dx_b
        const double syy = x1.y + y1.y;   // dy_a dy_b
        const double sA = x2.x + y2.x, sB = x2.y + y2.y;
        gdx[a] += sxx * dx[bb] + sab * dy[bb];
        gdx[bb] += sxx * dx[a] + sba * dy[a];
        gdy[a] += sba * dx[bb] + syy * dy[bb];
        gdy[bb] += sab * dx[a] + syy * dy[a];
        gal[a] += sA * al[bb];
        gal[bb] += sA * al[a];
        gbe[a] += sB * be[bb];
        gbe[bb] += sB * be[a];
      }
    }
    const double gs = sqrt(double(fem::ONE_THIRD_F32));
    const double rr = (p == 0 || p == 3) ? gs : -gs;
    const double ss = (p < 2) ? gs : -gs;
    const double q4 = 0.25;
    const double dr[4] = {q4 + q4 * ss, -(q4 + q4 * ss), -(q4 - q4 * ss), q4 - q4 * ss};
    const double ds[4] = {q4 + q4 * rr, q4 - q4 * rr, -(q4 - q4 * rr), -(q4 + q4 * rr)};
    const double fa[4] = {1.0 + ss, 1.0 + ss, 1.0 - ss, 1.0 - ss}, fb[4] = {1.0 + rr, 1.0 - rr, 1.0 - rr, 1.0 + rr};
    double s7[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      s7[0] += ((dx[a] * gdx[a] + dy[a] * gdy[a]) + al[a] * gal[a]) + be[a] * gbe[a];
      s7[1] += gdx[a] * dr[a];
      s7[2] += gdx[a] * ds[a];
      s7[3] += gdy[a] * dr[a];
      s7[4] += gdy[a] * ds[a];
      s7[5] += gal[a] * fa[a];
      s7[6] += gbe[a] * fb[a];
    }
    s7[0] *= 0.5;
#pragma unroll
    for (int k = 1; k < 7; ++k) s7[k] *= w;
    fem::plate_point_reverse_r(xl, p, s7, rdet, rqr, rqs, gxl);
  }
  // the parameters again (not held through the sweeps)
  plate_write_flat(A, e, p, live, A.params[4 * e], A.params[4 * e + 1], A.params[4 * e + 2], A.params[4 * e + 3], P,
                   gxl);
}

// Compliance sensitivities of plates by strain energy: four lanes per element,
// lane p = integration point p (and node p's drilling term and outputs).  Lane
// p loads node p's displacements; a quad broadcast gives every lane all four.
template <bool FLAT>
__global__ __launch_bounds__(256) void plate_compliance_kernel(AdjArgs A) {
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t e0 = size_t(b) * 64 + (threadIdx.x >> 2);
  const int p = int(threadIdx.x & 3u);
  const bool live = e0 < A.n;
  const size_t e = live ? e0 : A.n - 1;  // dead lanes still take part in the lane sums
  PlateElem el;
  load_plate_elem<FLAT>(A, e, el);
  const fem::PlateMat mat = fem::plate_material(el.E, el.nu, el.t, el.k);
  fem::PlateShear sv;
  fem::plate_shear_vectors(el.xl, sv);
  const uint32_t ap = p == 0 ? el.nd[0] : (p == 1 ? el.nd[1] : (p == 2 ? el.nd[2] : el.nd[3]));
  double up[6];
  if (FLAT) {
#pragma unroll
    for (int i = 0; i < 6; ++i) up[i] = A.g[6 * size_t(ap) + i];
  } else {
    local_u(A.g, ap, el.R, up);
  }
  double U[4][6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    U[0][i] = quad_bcast<0>(up[i]);
    U[1][i] = quad_bcast<1>(up[i]);
    U[2][i] = quad_bcast<2>(up[i]);
    U[3][i] = quad_bcast<3>(up[i]);
  }
  double P[7], gxl[6], hx = 0.0, hy = 0.0;
#pragma unroll
  for (int i = 0; i < 7; ++i) P[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = 0.0;
  fem::plate_compliance_point(el.xl, p, mat, sv, U, P, gxl, hx, hy);
  plate_write(A, e, p, live, el, P, gxl, hx, hy);
}

// ---------------------------------------------------------------- trusses
template <int MODE>
__global__ __launch_bounds__(256) void truss_adjoint_kernel(AdjArgs A) {
  const size_t e = size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
  if (e >= A.n) return;
  const uint32_t nd[2] = {A.conn[2 * e], A.conn[2 * e + 1]};
  double d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] = A.xyz[3 * size_t(nd[1]) + i] - A.xyz[3 * size_t(nd[0]) + i];
  const double r0[3] = {A.R[9 * e], A.R[9 * e + 1], A.R[9 * e + 2]};
  double H[9];
  if (MODE == MODE_DLDK) {
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = 0.0;
#pragma unroll 1
    for (int pq = 0; pq < 4; ++pq) {
      const int p = pq >> 1, q = pq & 1;
      const uint32_t k = A.bid[4 * e + pq], a = nd[p];
      const uint64_t um = A.bmask[k];
      const double s = p == q ? 1.0 : -1.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const uint32_t mi = uint32_t(um >> (6 * i)) & 63u;
        const int64_t rs = A.row_ptr[6 * size_t(a) + i] + A.boff[6 * size_t(k) + i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
          if ((mi >> j) & 1u) H[i * 3 + j] += s * A.g[rs + __builtin_popcount(mi & ((1u << j) - 1u))];
      }
    }
  } else {
    double du[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) du[i] = A.g[6 * size_t(nd[1]) + i] - A.g[6 * size_t(nd[0]) + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) H[i * 3 + j] = -(du[i] * du[j]);
  }
  double gp[3], gd[3];
  fem::truss_adjoint(d, r0, A.params[3 * e], A.params[3 * e + 1], A.params[3 * e + 2], H, gp, gd);
  if (A.dparam) {
#pragma unroll
    for (int i = 0; i < 3; ++i) A.dparam[3 * e + i] = gp[i];
  }
  if (A.xs) {
    double* x0 = A.xs + 3 * size_t(A.xpos[2 * e]);
    double* x1 = A.xs + 3 * size_t(A.xpos[2 * e + 1]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x0[i] = -gd[i];
      x1[i] = gd[i];
    }
  }
}

// ---------------------------------------------------------------- beams

// dL/dK: four lanes per beam, lane pq = block (p, q), so the four blocks'
// scattered gathers run in parallel; lane 0 adds the term sums in block order
// (pq = 0, 1, 2, 3) and writes the gradients.  Compliance: one lane per beam
// (its two nodal displacements serve all four blocks).
template <int MODE>
constexpr int beam_lanes() { return MODE == MODE_DLDK ? 4 : 1; }

template <int MODE>
__global__ __launch_bounds__(256) void beam_adjoint_kernel(AdjArgs A) {
  constexpr int LANES = beam_lanes<MODE>();
  const size_t e = (size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x) / LANES;
  const int pq0 = LANES == 4 ? int(threadIdx.x & 3u) : 0;
  if (e >= A.n) return;  // whole quads: the four lanes share e
  const uint32_t nd[2] = {A.conn[2 * e], A.conn[2 * e + 1]};
  double d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] = A.xyz[3 * size_t(nd[1]) + i] - A.xyz[3 * size_t(nd[0]) + i];
  const double L = fem::norm3(d[0], d[1], d[2]);
  const double J = fem::bar_jacobian(L);
  const double inv_j = 1.0 / J;
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = A.R[9 * e + k];
  const double* prm = A.params + 11 * e;
  double S[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int pq = pq0; pq < pq0 + 4 / LANES; ++pq) {
    const int p = pq >> 1, q = pq & 1;
    double gl[36];
    if (MODE == MODE_DLDK) {
      gather_block(A, A.bid[4 * e + pq], nd[p], gl);
      to_local(gl, R);
    } else {
      double up[6], uq[6];
      local_u(A.g, nd[p], R, up);
      local_u(A.g, nd[q], R, uq);
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) gl[i * 6 + j] = -(up[i] * uq[j]);
    }
    fem::beam_term_sums(gl, (p == 0 ? 0.0 - 0.5 : 0.0 + 0.5) * inv_j, (q == 0 ? 0.0 - 0.5 : 0.0 + 0.5) * inv_j, S);
  }
  if (LANES == 4) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double s1 = quad_bcast<1>(S[k]), s2 = quad_bcast<2>(S[k]), s3 = quad_bcast<3>(S[k]);
      S[k] = ((S[k] + s1) + s2) + s3;
    }
    if (pq0 != 0) return;
  }
  double i11p, i22p, angle;
  fem::beam_principal(prm[3], prm[4], prm[5], A.rel_tol, i11p, i22p, angle);
  double gp[11];
  fem::beam_adjoint_params(S, J * 2.0, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7],
                           A.derived[2 * e], A.derived[2 * e + 1], angle, gp);
  (void)i11p;
  (void)i22p;
  if (A.dparam) {
#pragma unroll
    for (int i = 0; i < 11; ++i) A.dparam[11 * e + i] = gp[i];
  }
}

// nodal coordinate gradient = sum of the node's (element, node) rows, which
// sit in consecutive slots in element-insertion order
__global__ __launch_bounds__(256) void node_gather_kernel(const uint32_t* __restrict__ inc_ptr,
                                                          const double* __restrict__ xs, size_t n_nodes,
                                                          double* __restrict__ dxyz) {
  const size_t a = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= n_nodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
    s0 += xs[3 * size_t(k)];
    s1 += xs[3 * size_t(k) + 1];
    s2 += xs[3 * size_t(k) + 2];
  }
  dxyz[3 * a] = s0;
  dxyz[3 * a + 1] = s1;
  dxyz[3 * a + 2] = s2;
}

}  // namespace

std::string upload_adjoint_maps(Device* d, const AdjointMaps& am) {
  for (int t = 0; t < 3; ++t) {
    auto e = dupload(d->adj_bid[t], am.bid[t].data(), am.bid[t].size());
    if (!e.empty()) return e;
  }
  auto e = dupload(d->adj_bmask, am.bmask.data(), am.bmask.size());
  if (!e.empty()) return e;
  e = dupload(d->adj_boff, am.boff.data(), am.boff.size());
  if (!e.empty()) return e;
  e = dupload(d->adj_inc_ptr, am.inc_ptr.data(), am.inc_ptr.size());
  if (!e.empty()) return e;
  e = dupload(d->adj_xpos, am.xpos.data(), am.xpos.size());
  if (!e.empty()) return e;
  for (int t = 0; t < 3; ++t) d->adj_xoff[t] = am.xoff[t];
  d->adj_flat = am.flat;
  if (am.flat) {
    e = dupload(d->adj_kpack, am.kpack.data(), am.kpack.size());
    if (!e.empty()) return e;
    e = dupload(d->adj_eorder, am.eorder.data(), am.eorder.size());
    if (!e.empty()) return e;
  }
  d->adj_nxs = am.n_xs_rows;
  return dalloc(d->adj_xs, size_t(3 * am.n_xs_rows));
}

std::string run_adjoint(Device* d, int mode, const double* g, double* const dparam[3], double* dxyz, void* stream,
                        bool wide) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipEventRecord(d->ev2, s));
  for (int t = 0; t < 3; ++t) {
    const size_t n = d->n_elem[t];
    if (!n) continue;
    AdjArgs A{};
    A.xyz = d->xyz;
    A.conn = d->conn[t];
    A.params = d->params[t];
    A.R = d->R[t];
    A.derived = d->derived;
    A.n = n;
    A.g = g;
    A.bid = d->adj_bid[t];
    A.bmask = d->adj_bmask;
    A.boff = d->adj_boff;
    A.kpack = reinterpret_cast<const uint2*>(d->adj_kpack);
    A.eorder = d->adj_eorder;
    A.row_ptr = d->row_ptr;
    A.dparam = dparam[t];
    A.xs = (dxyz && t != BEAM) ? d->adj_xs : nullptr;
    A.xpos = d->adj_xpos + d->adj_xoff[t];
    A.rel_tol = d->rel_tol;
    A.g_bytes = (mode == MODE_DLDK && wide && size_t(d->nnz) * 8 < (size_t(1) << 32))
                    ? uint32_t(size_t(d->nnz) * 8)
                    : 0u;
    if (t == PLATE) {
      const unsigned blocks = unsigned((n + 63) / 64);
      if (mode == MODE_DLDK && d->adj_flat)
        plate_adjoint_flat_kernel<<<unsigned((n + 15) / 16), 64, 0, s>>>(A);
      else if (mode == MODE_DLDK) plate_adjoint_kernel<<<blocks, 256, 0, s>>>(A);
      else if (d->adj_flat) plate_compliance_kernel<true><<<blocks, 256, 0, s>>>(A);
      else plate_compliance_kernel<false><<<blocks, 256, 0, s>>>(A);
    } else {
      const unsigned blocks = unsigned((n + 255) / 256);
      if (t == TRUSS) {
        if (mode == MODE_DLDK) truss_adjoint_kernel<MODE_DLDK><<<blocks, 256, 0, s>>>(A);
        else truss_adjoint_kernel<MODE_U><<<blocks, 256, 0, s>>>(A);
      } else {
        if (mode == MODE_DLDK)
          beam_adjoint_kernel<MODE_DLDK><<<unsigned((beam_lanes<MODE_DLDK>() * n + 255) / 256), 256, 0, s>>>(A);
        else
          beam_adjoint_kernel<MODE_U><<<unsigned((beam_lanes<MODE_U>() * n + 255) / 256), 256, 0, s>>>(A);
      }
    }
    HIP_TRY(hipGetLastError());
  }
  if (dxyz && d->n_nodes) {
    const unsigned blocks = unsigned((d->n_nodes + 255) / 256);
    node_gather_kernel<<<blocks, 256, 0, s>>>(d->adj_inc_ptr, d->adj_xs, d->n_nodes, dxyz);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(d->ev3, s));
  d->timed_bwd = true;
  return "";
}

double last_backward_ms(Device* d) {
  if (!d->timed_bwd) return -1.0;
  if (hipEventSynchronize(d->ev3) != hipSuccess) return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, d->ev2, d->ev3) != hipSuccess) return -1.0;
  return double(ms);
}

}  // namespace femdev
