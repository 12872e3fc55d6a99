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
// per-(element, node) rows summed per node in element-insertion order by
// node_gather_kernel.  No atomics.
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
  const int64_t* row_ptr;
  double* dparam;  // n x params_per, may be null
  double* xs;      // n x nodes_per x 3, may be null
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
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
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

__device__ inline double quad_sum(double v) {
  v += __shfl_xor(v, 1, 4);
  v += __shfl_xor(v, 2, 4);
  return v;
}

// ---------------------------------------------------------------- plates
// Four lanes per element (lane p = row block of node p), 64 elements per block.
// Flat standard block (13 values, 14 on the diagonal) held compactly:
// (u,u) (u,v) (v,u) (v,v) | (w,tx,ty) x (w,tx,ty) row-major | (tz,tz)
struct FlatBlock {
  double v[14];
};

__device__ inline void load_flat(const AdjArgs& A, const int64_t rp[6], const uint32_t bo[6], bool self,
                                 FlatBlock& g) {
  g.v[0] = A.g[rp[0] + bo[0]];
  g.v[1] = A.g[rp[0] + bo[0] + 1];
  g.v[2] = A.g[rp[1] + bo[1]];
  g.v[3] = A.g[rp[1] + bo[1] + 1];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) g.v[4 + 3 * i + j] = A.g[rp[2 + i] + bo[2 + i] + j];
  g.v[13] = self ? A.g[rp[5] + bo[5]] : 0.0;
}

// block k of a node whose nb blocks' values start at gr (fixed flat layout)
__device__ inline void load_flat_k(const double* gr, int nb, int k, bool self, FlatBlock& g) {
  g.v[0] = gr[2 * k];
  g.v[1] = gr[2 * k + 1];
  g.v[2] = gr[2 * nb + 2 * k];
  g.v[3] = gr[2 * nb + 2 * k + 1];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) g.v[4 + 3 * i + j] = gr[(4 + 3 * i) * nb + 3 * k + j];
  g.v[13] = self ? gr[13 * nb] : 0.0;
}

__device__ inline void expand_flat(const FlatBlock& g, double G[36]) {
#pragma unroll
  for (int i = 0; i < 36; ++i) G[i] = 0.0;
  G[0] = g.v[0];
  G[1] = g.v[1];
  G[6] = g.v[2];
  G[7] = g.v[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[(2 + i) * 6 + 2 + j] = g.v[4 + 3 * i + j];
  G[35] = g.v[13];
}

__device__ inline void load_boff(const AdjArgs& A, uint32_t k, uint32_t bo[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) bo[i] = A.boff[6 * size_t(k) + i];
}

// FLAT: every plate frame is the identity.  Every dL/dK block is then
// standard: the frame-rotation term vanishes identically (the pattern has no
// in-plane / out-of-plane couplings) and the blocks have fixed layouts.
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
#pragma unroll
    for (int i = 0; i < 3; ++i) A.xs[(4 * e + p) * 3 + i] = fem::sel4(gX[0][i], gX[1][i], gX[2][i], gX[3][i], p);
  }
}

template <bool FLAT>
__global__ __launch_bounds__(256) void plate_adjoint_kernel(AdjArgs A) {
  __shared__ double2 pts[64 * 4 * PA_POINT];
  const unsigned b = xcd_block(blockIdx.x, gridDim.x);
  const size_t e0 = size_t(b) * 64 + (threadIdx.x >> 2);
  const int p = int(threadIdx.x & 3u);
  const bool live = e0 < A.n;
  const size_t e = live ? e0 : A.n - 1;  // dead lanes still take part in the lane sums
  PlateElem el;
  load_plate_elem<FLAT>(A, e, el);
  const uint32_t* nd = el.nd;
  const double* R = el.R;
  const bool ident = FLAT || is_identity(R);
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
  // flat dL/dK blocks: the 13 (14 on the diagonal) values of block k of node
  // ap sit at fixed offsets from its first row (2 per row in the two membrane
  // rows, 3 per row in the three bending rows, then the drilling entry).  The
  // row position comes with the block indices, so the value loads wait for
  // one load only; each block's values are loaded while the previous block is
  // processed.
  uint32_t g0 = 0, kp = 0;
  int nb = 0;
  FlatBlock nxt;
  if (FLAT) {
    const uint2 kr = A.kpack[4 * e + p];
    g0 = kr.x;
    kp = kr.y;
    nb = int(kp >> 24);
    load_flat_k(A.g + g0, nb, int(kp & 63u), p == 0, nxt);
  }
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
    if (FLAT) {
      const FlatBlock g = nxt;
      if (q < 3) load_flat_k(A.g + g0, nb, int((kp >> (6 * (q + 1))) & 63u), p == q + 1, nxt);
      expand_flat(g, gl);
    } else {
      gather_block(A, A.bid[16 * e + 4 * p + q], ap, gl);
      if (!ident) to_local(gl, R);
    }
    double Mq[6], vq[3], uq_[3], cq[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Mq[k] = fem::sel4(acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3], q);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      vq[i] = fem::sel4(sv.v[0][i], sv.v[1][i], sv.v[2][i], sv.v[3][i], q);
      uq_[i] = fem::sel4(sv.u[0][i], sv.u[1][i], sv.u[2][i], sv.u[3][i], q);
    }
    // flat dL/dK blocks carry no in-plane / out-of-plane coupling: no frame term
    fem::plate_block_adjoint_q(Mq, mat, vp_, up_, vq, uq_, p, q, gl, P, cq, gxl, hx, hy, !FLAT);
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
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      A.xs[(2 * e) * 3 + i] = -gd[i];
      A.xs[(2 * e + 1) * 3 + i] = gd[i];
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

// nodal coordinate gradient = sum of the node's (element, node) rows in
// element-insertion order
__global__ __launch_bounds__(256) void node_gather_kernel(const uint32_t* __restrict__ inc_ptr,
                                                          const uint32_t* __restrict__ inc,
                                                          const double* __restrict__ xs, size_t n_nodes,
                                                          double* __restrict__ dxyz) {
  const size_t a = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= n_nodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
    const size_t r = inc[k];
    s0 += xs[3 * r];
    s1 += xs[3 * r + 1];
    s2 += xs[3 * r + 2];
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
  e = dupload(d->adj_inc, am.inc.data(), am.inc.size());
  if (!e.empty()) return e;
  for (int t = 0; t < 3; ++t) d->adj_xoff[t] = am.xoff[t];
  d->adj_flat = am.flat;
  if (am.flat) {
    e = dupload(d->adj_kpack, am.kpack.data(), am.kpack.size());
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
    A.row_ptr = d->row_ptr;
    A.dparam = dparam[t];
    A.xs = (dxyz && t != BEAM) ? d->adj_xs + 3 * d->adj_xoff[t] : nullptr;
    A.rel_tol = d->rel_tol;
    A.g_bytes = (mode == MODE_DLDK && wide && size_t(d->nnz) * 8 < (size_t(1) << 32))
                    ? uint32_t(size_t(d->nnz) * 8)
                    : 0u;
    if (t == PLATE) {
      const unsigned blocks = unsigned((n + 63) / 64);
      if (mode == MODE_DLDK && d->adj_flat) plate_adjoint_kernel<true><<<blocks, 256, 0, s>>>(A);
      else if (mode == MODE_DLDK) plate_adjoint_kernel<false><<<blocks, 256, 0, s>>>(A);
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
    node_gather_kernel<<<blocks, 256, 0, s>>>(d->adj_inc_ptr, d->adj_inc, d->adj_xs, d->n_nodes, dxyz);
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
