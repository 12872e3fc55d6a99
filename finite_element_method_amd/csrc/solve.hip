// solve.hip -- the crate's sparse global analysis on the device
// (methods_for_global_analysis.rs:28-133, 135-275):
//   b   = r_a - K_ab u_b                                  (find_b_sparse)
//   K_aa u_a = b by preconditioned conjugate gradients, Jacobi or node-block
//   Jacobi                       (pcg_jacobi_csr / pcg_block_jacobi_csr)
//   r_r = (K_ba u_a + K_bb u_b) - r_b                     (find_r_r_sparse)
// on the matrices of the last separation (separate.hip).
//
// PCG is Saad's Algorithm 9.1 with x0 = 0, stopped when
// ||r||_2 <= max(rel_tol ||b||_2, abs_tol) and counting the x updates.  All
// reductions use a fixed grid and fixed trees, so a solve is bitwise
// reproducible.  The scalars (alpha, beta, r.z) stay on the device; the host
// reads r.r once per iteration for the stopping test.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "device_state.hpp"
#include "fem_device.hpp"

namespace femdev {

namespace {

constexpr int SV_B = 256;         // threads per block of the vector kernels
constexpr int SV_PARTS = 1024;    // blocks of the reductions (fixed: reproducible sums)
constexpr int SPMV_LANES = 8;     // lanes per CSR row
constexpr int BJ_MAX = 6;         // dofs per node block

// scalar slots (device): 0 r.z, 1 p.Ap, 2 alpha, 3 beta, 4 r.r, 5 (r.z)_new, 6 b.b, 7 status,
// 8 updates done, 9 stop (converged or capped: every later kernel of the loop
// returns at once), 10 converged
enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_RZN, S_BB, S_STATUS, S_ITER, S_DONE, S_CONV, S_COUNT };
constexpr int BJ_PER_WAVE = 10;   // node blocks per wave in the block-Jacobi kernels (6 lanes each)

template <class T>
std::string dgrow(T*& p, size_t& cap, size_t n) {
  if (n <= cap && p) return "";
  auto e = dalloc(p, n ? n : 1);
  if (!e.empty()) { cap = 0; return e; }
  cap = n ? n : 1;
  return "";
}

// y[r] = (base ? base[r] : 0) + sign * (A x)[r] - (sub ? sub[r] : 0); SPMV_LANES lanes per row,
// each summing its strided entries in order, combined by a fixed butterfly.
__global__ void sv_spmv(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, const double* __restrict__ v,
                        const double* __restrict__ x, const double* base, double sign, const double* __restrict__ sub,
                        double* y, int64_t n) {
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SPMV_LANES;
  const int lig = threadIdx.x & (SPMV_LANES - 1);
  if (r >= n) return;
  double acc = 0.0;
  for (int64_t e = rp[r] + lig; e < rp[r + 1]; e += SPMV_LANES) acc += v[e] * x[ci[e]];
#pragma unroll
  for (int o = SPMV_LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, SPMV_LANES);
  if (lig == 0) {
    double out = base ? base[r] + sign * acc : sign * acc;
    if (sub) out -= sub[r];
    y[r] = out;
  }
}

// Ap = K p inside the update loop (returns at once after the stop flag).
// SPMV_QUAD_LANES lanes per row, each taking four neighbouring values per step
// with two 16-byte value loads and one 16-byte column load, range-checked
// buffer loads (32-bit offsets: below 4 GiB of values).  On the plate matrices
// (rows of 9 to 27 values) this beats one 8-byte load per lane and value by
// about a fifth: fewer, fuller memory requests.  Fixed summation order (each
// lane's values in turn, then a butterfly over the row's lanes).
constexpr int SPMV_QUAD_LANES = 4;
__global__ void sv_spmv_it_quad(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                const double* __restrict__ v, const double* __restrict__ p, double* __restrict__ ap,
                                int64_t n, const double* s, uint32_t v_bytes) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  if (s[S_DONE] != 0.0) return;
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SPMV_QUAD_LANES;
  const int lig = threadIdx.x & (SPMV_QUAD_LANES - 1);
  if (r >= n) return;
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(v), 0, int(v_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(ci), 0, int(v_bytes / 2u), 0x00020000);
  const int64_t end = rp[r + 1];
  double acc = 0.0;
  for (int64_t e = rp[r] + 4 * lig; e < end; e += 4 * SPMV_QUAD_LANES) {
    const uint32_t ev = uint32_t(e);
    // values and columns past the row end (or the array end: read as 0) are
    // loaded but not used; their columns are valid indices or 0
    const double2 w0 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rv, 8u * ev, 0, 0));
    const double2 w1 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rv, 8u * ev + 16u, 0, 0));
    const u32x4 c = __builtin_amdgcn_raw_buffer_load_b128(rc, 4u * ev, 0, 0);
    acc = __builtin_fma(w0.x, p[int32_t(c.x)], acc);
    if (e + 1 < end) acc = __builtin_fma(w0.y, p[int32_t(c.y)], acc);
    if (e + 2 < end) acc = __builtin_fma(w1.x, p[int32_t(c.z)], acc);
    if (e + 3 < end) acc = __builtin_fma(w1.y, p[int32_t(c.w)], acc);
  }
#pragma unroll
  for (int o = SPMV_QUAD_LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, SPMV_QUAD_LANES);
  if (lig == 0) ap[r] = acc;
}

// the same with one 8-byte load per lane and value (values beyond 4 GiB)
__global__ void sv_spmv_it(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ p, double* __restrict__ ap,
                           int64_t n, const double* s) {
  if (s[S_DONE] != 0.0) return;
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SPMV_LANES;
  const int lig = threadIdx.x & (SPMV_LANES - 1);
  if (r >= n) return;
  double acc = 0.0;
  for (int64_t e = rp[r] + lig; e < rp[r + 1]; e += SPMV_LANES) acc = __builtin_fma(v[e], p[ci[e]], acc);
#pragma unroll
  for (int o = SPMV_LANES / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, SPMV_LANES);
  if (lig == 0) ap[r] = acc;
}

// block tree sum of per-thread values into part[blockIdx.x]
__device__ inline void block_sum2(double a, double b, double* part_a, double* part_b) {
  __shared__ double sa[SV_B], sb[SV_B];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = SV_B / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = sa[0];
    if (part_b) part_b[blockIdx.x] = sb[0];
  }
}

// partials of p.Ap; grid SV_PARTS x SV_B, strided
__global__ void sv_dot_it(const double* __restrict__ x, const double* __restrict__ y, int64_t n, const double* s,
                          double* part) {
  if (s[S_DONE] != 0.0) return;
  double a = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * SV_B + threadIdx.x; i < n; i += int64_t(SV_PARTS) * SV_B)
    a = __builtin_fma(x[i], y[i], a);
  block_sum2(a, 0.0, part, nullptr);
}

// final sums of the partials into scalars, then the PCG scalar step:
// mode 0: s[S_RZ] = a, s[S_BB] = b (start);  mode 1: s[S_PAP] = a, alpha = rz / pAp;
// mode 2: s[S_RZN] = a, s[S_RR] = b, beta = rzn / rz, rz = rzn
__global__ void sv_final(const double* __restrict__ part_a, const double* __restrict__ part_b, double* s, int mode,
                         double stop, double max_iter) {
  if (mode != 0 && s[S_DONE] != 0.0) return;
  __shared__ double sa[SV_PARTS], sb[SV_PARTS];
  for (int i = threadIdx.x; i < SV_PARTS; i += blockDim.x) {
    sa[i] = part_a[i];
    sb[i] = part_b ? part_b[i] : 0.0;
  }
  __syncthreads();
  for (int st = SV_PARTS / 2; st > 0; st >>= 1) {
    for (int i = threadIdx.x; i < st; i += blockDim.x) {
      sa[i] += sa[i + st];
      sb[i] += sb[i + st];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    s[S_RZ] = sa[0];
    s[S_BB] = sb[0];
    s[S_RR] = sb[0];
  } else if (mode == 1) {
    s[S_PAP] = sa[0];
    s[S_ALPHA] = s[S_RZ] / sa[0];
  } else {
    s[S_RZN] = sa[0];
    s[S_RR] = sb[0];
    s[S_BETA] = sa[0] / s[S_RZ];
    s[S_RZ] = sa[0];
    const double it = s[S_ITER] + 1.0;
    const bool conv = sqrt(sb[0]) <= stop;
    s[S_ITER] = it;
    s[S_CONV] = conv ? 1.0 : 0.0;
    s[S_DONE] = (conv || it >= max_iter) ? 1.0 : 0.0;
  }
}

// Jacobi: d[i] = K_aa(i, i)
__global__ void sv_diag(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, const double* __restrict__ v,
                        int64_t n, double* __restrict__ d) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double dd = 0.0;
  for (int64_t e = rp[r]; e < rp[r + 1]; ++e)
    if (ci[e] == r) dd = v[e];
  d[r] = dd;
}

// node blocks of K_aa: flag[i] = row i starts a node (methods_for_global_analysis.rs:135-160)
__global__ void sv_block_flags(const int64_t* __restrict__ aa, int64_t n, uint32_t* __restrict__ flag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  flag[i] = i < n && (i == 0 || aa[i] / 6 != aa[i - 1] / 6);
}
__global__ void sv_block_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ id, int64_t n,
                                int32_t* __restrict__ starts) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) starts[id[n]] = int32_t(n);  // id[n] = number of blocks
  else if (flag[i]) starts[id[i]] = int32_t(i);
}

// inverse of each node block (Gauss-Jordan, partial pivoting); status bit 1 on a singular block
__global__ __launch_bounds__(256) void sv_block_inverse(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                        const double* __restrict__ v,
                                                        const int32_t* __restrict__ starts, int64_t nblk,
                                                        double* __restrict__ inv, double* s) {
  // The augmented block [A | I] lives in registers: every loop below runs over
  // compile-time indices, so row swaps and column placement are selects (a
  // runtime-indexed array would sit in scratch memory).  A block of m < 6 dofs
  // is padded with the identity: its rows and columns past m never become
  // pivots or elimination targets of the first m columns, so the m x m inverse
  // is the one the unpadded elimination gives.
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int s0 = starts[b], m = starts[b + 1] - s0;
  double* o = inv + 36 * b;
  if (m > BJ_MAX) {  // a node has at most six dofs; guard the fixed-size arrays anyway
    for (int i = 0; i < 36; ++i) o[i] = 0.0;
    s[S_STATUS] = 1.0;
    return;
  }
  double a[BJ_MAX][2 * BJ_MAX];
#pragma unroll
  for (int i = 0; i < BJ_MAX; ++i) {
#pragma unroll
    for (int j = 0; j < 2 * BJ_MAX; ++j) a[i][j] = (j == BJ_MAX + i || (i >= m && j == i)) ? 1.0 : 0.0;
    if (i < m) {
      // the block's columns s0 .. s0 + m - 1 are one run of the sorted row:
      // find its start by bisection, then place at most m entries
      int64_t lo = rp[s0 + i], hi = rp[s0 + i + 1];
      const int64_t end = hi;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (ci[mid] < s0) lo = mid + 1; else hi = mid;
      }
      for (int64_t e = lo; e < end; ++e) {
        const int c = ci[e] - s0;
        if (c >= m) break;
        const double x = v[e];
#pragma unroll
        for (int j = 0; j < BJ_MAX; ++j) a[i][j] = c == j ? x : a[i][j];
      }
    }
  }
  bool ok = true;
#pragma unroll
  for (int c = 0; c < BJ_MAX; ++c) {
    // partial pivoting: the first row of largest |a[i][c]|, i >= c
    int piv = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int i = c + 1; i < BJ_MAX; ++i)
      if (fabs(a[i][c]) > best) {
        piv = i;
        best = fabs(a[i][c]);
      }
    ok = ok && best != 0.0;
#pragma unroll
    for (int i = c + 1; i < BJ_MAX; ++i) {
      const bool sw = piv == i;
#pragma unroll
      for (int j = 0; j < 2 * BJ_MAX; ++j) {
        const double t = a[c][j];
        a[c][j] = sw ? a[i][j] : t;
        a[i][j] = sw ? t : a[i][j];
      }
    }
    const double d = a[c][c];
#pragma unroll
    for (int j = 0; j < 2 * BJ_MAX; ++j) a[c][j] /= d;
#pragma unroll
    for (int i = 0; i < BJ_MAX; ++i)
      if (i != c) {
        const double f = a[i][c];
#pragma unroll
        for (int j = 0; j < 2 * BJ_MAX; ++j) a[i][j] = f != 0.0 ? a[i][j] - f * a[c][j] : a[i][j];
      }
  }
#pragma unroll
  for (int i = 0; i < BJ_MAX; ++i)
#pragma unroll
    for (int j = 0; j < BJ_MAX; ++j) o[6 * i + j] = (i < m && j < m) ? a[i][BJ_MAX + j] : 0.0;
  if (!ok) s[S_STATUS] = 1.0;  // any writer stores the same value
}

// start: r = b, z = M^-1 r, p = z, partials of (r, z) and (b, b)
__global__ void sv_start_jacobi(const double* __restrict__ b, const double* __restrict__ d, int64_t n,
                                double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                double* __restrict__ p, double* part_a, double* part_b) {
  double ra = 0.0, rb = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * SV_B + threadIdx.x; i < n; i += int64_t(SV_PARTS) * SV_B) {
    const double bi = b[i], zi = bi / d[i];
    x[i] = 0.0;
    r[i] = bi;
    z[i] = zi;
    p[i] = zi;
    ra += bi * zi;
    rb += bi * bi;
  }
  block_sum2(ra, rb, part_a, part_b);
}

// update: x += alpha p, r -= alpha Ap, z = r / d; partials of (r, z) and (r, r)
__global__ void sv_update_jacobi(const double* __restrict__ d, const double* __restrict__ p,
                                 const double* __restrict__ ap, const double* __restrict__ s, int64_t n,
                                 double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, double* part_a,
                                 double* part_b) {
  if (s[S_DONE] != 0.0) return;
  const double alpha = s[S_ALPHA];
  double ra = 0.0, rb = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * SV_B + threadIdx.x; i < n; i += int64_t(SV_PARTS) * SV_B) {
    x[i] = __builtin_fma(alpha, p[i], x[i]);
    const double ri = __builtin_fma(-alpha, ap[i], r[i]);
    r[i] = ri;
    const double zi = ri / d[i];
    z[i] = zi;
    ra += ri * zi;
    rb += ri * ri;
  }
  block_sum2(ra, rb, part_a, part_b);
}

// block Jacobi: six lanes per node block (lane = row of the block), ten
// blocks per wave; z = M^-1 r with the row of the inverse read as three
// 16-byte loads and the block's r gathered by lane shuffles.  Lanes beyond
// a block's size, and the last four lanes of each wave, only take part in the
// shuffles.
struct BjLane {
  int64_t k;  // node block
  int row, s0, m;
  bool act;   // row < m of an existing block
};
__device__ inline BjLane bj_lane(const int32_t* __restrict__ starts, int64_t nblk, int64_t wave, int lane) {
  BjLane o;
  o.k = wave * BJ_PER_WAVE + lane / 6;
  o.row = lane % 6;
  o.s0 = 0;
  o.m = 0;
  if (lane < 6 * BJ_PER_WAVE && o.k < nblk) {
    o.s0 = starts[o.k];
    o.m = starts[o.k + 1] - o.s0;
  }
  o.act = o.row < o.m;
  return o;
}
__device__ inline double bj_row_apply(const double* __restrict__ inv, const BjLane& L, int lane, double ri) {
  double2 w0 = make_double2(0.0, 0.0), w1 = w0, w2 = w0;
  if (L.act) {
    const double2* row = reinterpret_cast<const double2*>(inv + 36 * L.k + 6 * L.row);
    w0 = row[0];
    w1 = row[1];
    w2 = row[2];
  }
  const double w[6] = {w0.x, w0.y, w1.x, w1.y, w2.x, w2.y};
  double zi = 0.0;
  const int first = lane - L.row;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const double rj = __shfl(ri, (first + j) & 63);  // every lane takes part
    if (j < L.m) zi = __builtin_fma(w[j], rj, zi);
  }
  return zi;
}
__global__ void sv_start_block(const double* __restrict__ b, const double* __restrict__ inv,
                               const int32_t* __restrict__ starts, int64_t nblk, double* __restrict__ x,
                               double* __restrict__ r, double* __restrict__ z, double* __restrict__ p, double* part_a,
                               double* part_b) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = int64_t(SV_PARTS) * (SV_B / 64);
  double ra = 0.0, rb = 0.0;
  for (int64_t w = int64_t(blockIdx.x) * (SV_B / 64) + (threadIdx.x >> 6); w * BJ_PER_WAVE < nblk; w += waves) {
    const BjLane L = bj_lane(starts, nblk, w, lane);
    double ri = 0.0;
    if (L.act) {
      const int64_t i = L.s0 + L.row;
      ri = b[i];
      x[i] = 0.0;
      r[i] = ri;
    }
    const double zi = bj_row_apply(inv, L, lane, ri);
    if (L.act) {
      const int64_t i = L.s0 + L.row;
      z[i] = zi;
      p[i] = zi;
      ra = __builtin_fma(ri, zi, ra);
      rb = __builtin_fma(ri, ri, rb);
    }
  }
  block_sum2(ra, rb, part_a, part_b);
}
// update: x += alpha p, r -= alpha Ap, z = M^-1 r; partials of (r, z) and (r, r)
__global__ void sv_update_block(const double* __restrict__ inv, const int32_t* __restrict__ starts, int64_t nblk,
                                const double* __restrict__ p, const double* __restrict__ ap,
                                const double* __restrict__ s, double* __restrict__ x, double* __restrict__ r,
                                double* __restrict__ z, double* part_a, double* part_b) {
  if (s[S_DONE] != 0.0) return;
  const double alpha = s[S_ALPHA];
  const int lane = threadIdx.x & 63;
  const int64_t waves = int64_t(SV_PARTS) * (SV_B / 64);
  double ra = 0.0, rb = 0.0;
  for (int64_t w = int64_t(blockIdx.x) * (SV_B / 64) + (threadIdx.x >> 6); w * BJ_PER_WAVE < nblk; w += waves) {
    const BjLane L = bj_lane(starts, nblk, w, lane);
    double ri = 0.0;
    if (L.act) {
      const int64_t i = L.s0 + L.row;
      x[i] = __builtin_fma(alpha, p[i], x[i]);
      ri = __builtin_fma(-alpha, ap[i], r[i]);
      r[i] = ri;
    }
    const double zi = bj_row_apply(inv, L, lane, ri);
    if (L.act) {
      z[L.s0 + L.row] = zi;
      ra = __builtin_fma(ri, zi, ra);
      rb = __builtin_fma(ri, ri, rb);
    }
  }
  block_sum2(ra, rb, part_a, part_b);
}

// p = z + beta p
__global__ void sv_direction(const double* __restrict__ z, const double* __restrict__ s, int64_t n,
                             double* __restrict__ p) {
  if (s[S_DONE] != 0.0) return;
  const double beta = s[S_BETA];
  for (int64_t i = int64_t(blockIdx.x) * SV_B + threadIdx.x; i < n; i += int64_t(SV_PARTS) * SV_B)
    p[i] = __builtin_fma(beta, p[i], z[i]);
}

// r_b: forces at the k_bb indexes
__global__ void sv_gather(const double* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                          double* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

unsigned grid_for(int64_t items, int per_block) { return unsigned(std::max<int64_t>(1, (items + per_block - 1) / per_block)); }

}  // namespace

// Work buffers of one solve.  get(slot, bytes) returns device memory for a
// slot; vectors: 0 b, 1 x = u_a, 2 r, 3 z, 4 p, 5 Ap, 6 r_r, 7 r_b; 8 the
// reduction partials, 9 the scalars, 10 the Jacobi diagonal, 11 block flags,
// 12 block ids, 13 block starts, 14 scan storage, 15 block inverses.
struct SolveBuffers {
  virtual std::string get(int slot, size_t bytes, void** out) = 0;
  double* host = nullptr;  // S_COUNT doubles of host memory for the scalar reads
};

namespace {

__global__ void sv_zero(double* __restrict__ x, int n) {
  if (int(threadIdx.x) < n) x[threadIdx.x] = 0.0;
}

template <class T>
std::string take(SolveBuffers& bufs, int slot, size_t n, T*& out) {
  void* p = nullptr;
  auto e = bufs.get(slot, std::max<size_t>(n, 1) * sizeof(T), &p);
  out = static_cast<T*>(p);
  return e;
}

// the solve proper: b, preconditioner, PCG updates, r_r, into `bufs`
std::string solve_core(Device* d, int precond, int max_iter, double rel_tol, double abs_tol, hipStream_t s,
                       SolveResult& res, bool wide, SolveBuffers& bufs) {
  res = SolveResult();
  const int64_t na = d->sep_rows[0], nb = d->sep_rows[2];
  const size_t una = size_t(std::max<int64_t>(na, 1)), unb = size_t(std::max<int64_t>(nb, 1));
  std::string e;
  double* vec[8];
  for (int k = 0; k < 6; ++k)  // b, x, r, z, p, Ap
    if (!(e = take(bufs, k, una, vec[k])).empty()) return e;
  if (!(e = take(bufs, 6, unb, vec[6])).empty()) return e;
  if (!(e = take(bufs, 7, unb, vec[7])).empty()) return e;
  double *part = nullptr, *sc = nullptr;
  if (!(e = take(bufs, 8, size_t(2 * SV_PARTS), part)).empty()) return e;
  if (!(e = take(bufs, 9, size_t(S_COUNT), sc)).empty()) return e;
  double *b = vec[0], *x = vec[1], *r = vec[2], *z = vec[3], *p = vec[4], *ap = vec[5], *rr = vec[6], *rb = vec[7];
  double *pa = part, *pb = part + SV_PARTS;
  double* host = bufs.host;
  const int64_t* rp[4];
  const int32_t* ci[4];
  const double* v[4];
  for (int w = 0; w < 4; ++w)
    if (!(e = separated_csr(d, w, &rp[w], &ci[w], &v[w])).empty()) return e;
  const int64_t *aa_idx = d->sep_aa, *bb_idx = d->sep_bb;
  const double *r_a = d->sep_ra, *u_b = d->sep_ub;
  const unsigned B = SV_B;
  HIP_TRY(hipEventRecord(d->ev4, s));
  // b = r_a - K_ab u_b
  if (na) {
    sv_spmv<<<grid_for(na * SPMV_LANES, B), B, 0, s>>>(rp[1], ci[1], v[1], u_b, r_a, -1.0, nullptr, b, na);
    HIP_TRY(hipGetLastError());
  }
  // preconditioner
  int64_t nblk = 0;
  double* diag = nullptr;
  double* inv = nullptr;
  int32_t* starts = nullptr;
  sv_zero<<<1, 64, 0, s>>>(sc, S_COUNT);
  HIP_TRY(hipGetLastError());
  if (precond == 0) {
    if (!(e = take(bufs, 10, una, diag)).empty()) return e;
    if (na) {
      sv_diag<<<grid_for(na, B), B, 0, s>>>(rp[0], ci[0], v[0], na, diag);
      HIP_TRY(hipGetLastError());
    }
  } else {
    uint32_t *flag = nullptr, *bid = nullptr;
    if (!(e = take(bufs, 11, una + 1, flag)).empty()) return e;
    if (!(e = take(bufs, 12, una + 1, bid)).empty()) return e;
    if (!(e = take(bufs, 13, una + 1, starts)).empty()) return e;
    sv_block_flags<<<grid_for(na + 1, B), B, 0, s>>>(aa_idx, na, flag);
    HIP_TRY(hipGetLastError());
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag, bid, int(na + 1), s));
    void* tmp = nullptr;
    if (!(e = bufs.get(14, std::max<size_t>(bytes, 1), &tmp)).empty()) return e;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, flag, bid, int(na + 1), s));
    sv_block_starts<<<grid_for(na + 1, B), B, 0, s>>>(flag, bid, na, starts);
    HIP_TRY(hipGetLastError());
    uint32_t nb32 = 0;
    HIP_TRY(hipMemcpyAsync(host, bid + na, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::memcpy(&nb32, host, sizeof(uint32_t));
    nblk = int64_t(nb32);
    if (!(e = take(bufs, 15, size_t(36 * std::max<int64_t>(nblk, 1)), inv)).empty()) return e;
    if (nblk) {
      sv_block_inverse<<<grid_for(nblk, B), B, 0, s>>>(rp[0], ci[0], v[0], starts, nblk, inv, sc);
      HIP_TRY(hipGetLastError());
    }
  }
  // start
  if (precond == 0)
    sv_start_jacobi<<<SV_PARTS, B, 0, s>>>(b, diag, na, x, r, z, p, pa, pb);
  else
    sv_start_block<<<SV_PARTS, B, 0, s>>>(b, inv, starts, nblk, x, r, z, p, pa, pb);
  HIP_TRY(hipGetLastError());
  sv_final<<<1, 1024, 0, s>>>(pa, pb, sc, 0, 0.0, 0.0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(host, sc, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (host[S_STATUS] != 0.0) return "PCG(BlockJacobi): singular node block";
  const double stop = std::max(rel_tol * std::sqrt(host[S_BB]), abs_tol);
  int it = 0;
  bool conv = std::sqrt(host[S_RR]) <= stop;
  if (!conv && max_iter > 0) {
    // Updates are queued in growing batches; the last scalar kernel of an
    // update raises the stop flag on convergence or at max_iter, and every
    // kernel queued after that returns at once.  The count stays exact and the
    // host waits once per batch, not once per update.
    const bool quad = wide && size_t(d->sep_nnz[0]) * 8 < (size_t(1) << 32);
    const uint32_t v_bytes = quad ? uint32_t(size_t(d->sep_nnz[0]) * 8) : 0u;
    const unsigned spmv_blocks = grid_for(na * (quad ? SPMV_QUAD_LANES : SPMV_LANES), B);
    int queued = 0, batch = 1;
    for (;;) {
      const int nq = std::min(batch, max_iter - queued);
      for (int q = 0; q < nq; ++q) {
        if (quad) sv_spmv_it_quad<<<spmv_blocks, B, 0, s>>>(rp[0], ci[0], v[0], p, ap, na, sc, v_bytes);
        else sv_spmv_it<<<spmv_blocks, B, 0, s>>>(rp[0], ci[0], v[0], p, ap, na, sc);
        sv_dot_it<<<SV_PARTS, B, 0, s>>>(p, ap, na, sc, pa);
        sv_final<<<1, 1024, 0, s>>>(pa, nullptr, sc, 1, stop, double(max_iter));
        if (precond == 0)
          sv_update_jacobi<<<SV_PARTS, B, 0, s>>>(diag, p, ap, sc, na, x, r, z, pa, pb);
        else
          sv_update_block<<<SV_PARTS, B, 0, s>>>(inv, starts, nblk, p, ap, sc, x, r, z, pa, pb);
        sv_final<<<1, 1024, 0, s>>>(pa, pb, sc, 2, stop, double(max_iter));
        sv_direction<<<SV_PARTS, B, 0, s>>>(z, sc, na, p);
      }
      HIP_TRY(hipGetLastError());
      queued += nq;
      HIP_TRY(hipMemcpyAsync(host, sc, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (host[S_DONE] != 0.0 || queued >= max_iter) break;
      batch = std::min(batch * 2, 64);
    }
    it = int(host[S_ITER]);
    conv = host[S_CONV] != 0.0;
  }
  // r_r = (K_ba u_a + K_bb u_b) - r_b
  if (nb) {
    sv_gather<<<grid_for(nb, B), B, 0, s>>>(d->sep_force, bb_idx, nb, rb);
    sv_spmv<<<grid_for(nb * SPMV_LANES, B), B, 0, s>>>(rp[2], ci[2], v[2], x, nullptr, 1.0, nullptr, rr, nb);
    sv_spmv<<<grid_for(nb * SPMV_LANES, B), B, 0, s>>>(rp[3], ci[3], v[3], u_b, rr, 1.0, rb, rr, nb);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(d->ev5, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, d->ev4, d->ev5));
  d->sol_ms = ms;
  res.iterations = it;
  res.converged = conv;
  res.residual = std::sqrt(host[S_RR]);
  res.stop = stop;
  return "";
}

// the model's own buffers, grown on demand
struct ModelBuffers : SolveBuffers {
  Device* d;
  explicit ModelBuffers(Device* dev) : d(dev) {}
  std::string get(int slot, size_t bytes, void** out) override {
    std::string e;
    const size_t n8 = (bytes + 7) / 8, n4 = (bytes + 3) / 4;
    if (slot < 6) { e = dgrow(d->sol_vec[slot], d->sol_cap[slot], n8); *out = d->sol_vec[slot]; }
    else if (slot == 6) { e = dgrow(d->sol_rr, d->sol_cap[6], n8); *out = d->sol_rr; }
    else if (slot == 7) { e = dgrow(d->sol_rb, d->sol_cap[7], n8); *out = d->sol_rb; }
    else if (slot == 8) { e = dgrow(d->sol_part, d->sol_cap[8], n8); *out = d->sol_part; }
    else if (slot == 9) { e = dgrow(d->sol_s, d->sol_cap[9], n8); *out = d->sol_s; }
    else if (slot == 10) { e = dgrow(d->sol_diag, d->sol_cap[10], n8); *out = d->sol_diag; }
    else if (slot == 11) { e = dgrow(d->sol_flag, d->sol_cap[11], n4); *out = d->sol_flag; }
    else if (slot == 12) { e = dgrow(d->sol_bid, d->sol_cap[12], n4); *out = d->sol_bid; }
    else if (slot == 13) { e = dgrow(d->sol_starts, d->sol_cap[13], n4); *out = d->sol_starts; }
    else if (slot == 14) { e = dgrow(d->sol_tmp, d->sol_cap[14], bytes); *out = d->sol_tmp; }
    else { e = dgrow(d->sol_inv, d->sol_cap[15], n8); *out = d->sol_inv; }
    return e;
  }
};

}  // namespace

std::string solve_pcg(Device* d, int precond, int max_iter, double rel_tol, double abs_tol, void* stream,
                      SolveResult& res, bool wide) {
  d->sol_done = false;
  if (!d->sep_done) return "no separation computed";
  if (!d->sol_host) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&d->sol_host), 16 * sizeof(double)));
  ModelBuffers bufs(d);
  bufs.host = d->sol_host;
  auto e = solve_core(d, precond, max_iter, rel_tol, abs_tol, static_cast<hipStream_t>(stream), res, wide, bufs);
  if (!e.empty()) return e;
  d->sol_na = d->sep_rows[0];
  d->sol_nb = d->sep_rows[2];
  d->sol_done = true;
  return "";
}

// ---- the solve without allocation: every work buffer in one static array per GPU

namespace {

static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

// device address of `scratch` on each GPU, looked up once per device (the
// array's name is only a host-side stand-in for it)
std::string scratch_address(void** out) {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return "device index out of range for the static scratch table";
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return "hipGetSymbolAddress returned no address for the static scratch";
    addr[dev] = p;
  }
  *out = addr[dev];
  return "";
}

// slots 1 (u_a) and 6 (r_r) are the caller's outputs; the others are carved
// from the static array, each on a 256-byte boundary
struct StaticBuffers : SolveBuffers {
  unsigned char* base = nullptr;
  size_t used = 0;
  double *u_a = nullptr, *r_r = nullptr;
  std::string get(int slot, size_t bytes, void** out) override {
    if (slot == 1) { *out = u_a; return ""; }
    if (slot == 6) { *out = r_r; return ""; }
    const size_t off = (used + 255) & ~size_t(255);
    if (off + bytes > STATIC_SCRATCH_BYTES)
      return "the solve's work buffers exceed the static scratch of " + std::to_string(STATIC_SCRATCH_BYTES >> 20) +
             " MiB";
    *out = base + off;
    used = off + bytes;
    return "";
  }
};

}  // namespace

std::string solve_pcg_static(Device* d, int precond, int max_iter, double rel_tol, double abs_tol, void* stream,
                             double* u_a, double* r_r, SolveResult& res, bool wide) {
  if (!d->sep_done) return "no separation computed";
  if ((!u_a && d->sep_rows[0]) || (!r_r && d->sep_rows[2])) return "u_a and r_r must be device buffers";
  void* base = nullptr;
  auto e = scratch_address(&base);
  if (!e.empty()) return e;
  StaticBuffers bufs;
  bufs.base = static_cast<unsigned char*>(base);
  bufs.u_a = u_a;
  bufs.r_r = r_r;
  double host[S_COUNT];
  bufs.host = host;
  return solve_core(d, precond, max_iter, rel_tol, abs_tol, static_cast<hipStream_t>(stream), res, wide, bufs);
}

std::string solution(Device* d, const double** u_a, const double** r_r, int64_t* n_aa, int64_t* n_bb) {
  if (!d->sol_done) return "no solution computed";
  if (u_a) *u_a = d->sol_vec[1];
  if (r_r) *r_r = d->sol_rr;
  if (n_aa) *n_aa = d->sol_na;
  if (n_bb) *n_bb = d->sol_nb;
  return "";
}

double last_solve_ms(Device* d) { return d->sol_ms; }

}  // namespace femdev
