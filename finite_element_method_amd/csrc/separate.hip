// separate.hip -- sparse partition of the assembled stiffness into K_aa, K_ab,
// K_ba, K_bb with k_aa_indexes, k_bb_indexes, r_a and u_b
// (methods_for_separate_stiffness_matrix.rs:217-345) on the device.
//
// A mask-and-compact over the CSR with no atomics and one host read at the end:
//   classify  per row: stored diagonal (0 when absent), class (0 excluded,
//             1 free = a, 2 constrained = b), error key
//   scan      packed (a, b) row positions (one 64-bit exclusive sum)
//   count     per kept row: non-zero entries per target matrix, k_aa/k_bb
//             indexes, r_a and u_b
//   scan      all four matrices' row counts in one buffer
//   fill      per kept row: local columns and values, in column order
//   finish    per-matrix row pointers (rebased to 0) and the summary
// Rows and columns of every output come out sorted.  The crate returns the
// first constrained dof without stiffness in index order: that is the minimum
// error key.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <string>

#include "device_state.hpp"
#include "fem_device.hpp"

namespace femdev {

namespace {

constexpr int SEP_SUMMARY = 12;  // err, n_aa, n_bb, nnz[4], entry offset[4], spare

template <class T>
std::string dgrow(T*& p, size_t& cap, size_t n) {
  if (n <= cap && p) return "";
  auto e = dalloc(p, n ? n : 1);
  if (!e.empty()) { cap = 0; return e; }
  cap = n ? n : 1;
  return "";
}

__device__ inline uint32_t lo32(uint64_t x) { return uint32_t(x); }
__device__ inline uint32_t hi32(uint64_t x) { return uint32_t(x >> 32); }

// G lanes per row (G a power of two <= 32): the lanes of a group stride over
// the row's entries together, so loads are coalesced; counts are combined
// with lane shuffles and output slots with wave ballots.  Every lane of a
// group has the same row, so groups enter and leave loops as a unit.
template <int G>
__device__ inline int group_sum(int x) {
  for (int o = G / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, G);
  return x;
}

// Position of each row's diagonal entry in the CSR (-1 when the pattern has
// none); depends on the pattern only, so it is computed once per pattern.
__global__ void sep_diag_pos(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t n,
                             int64_t* __restrict__ dpos) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t lo = rp[r], hi = rp[r + 1], found = -1;
  while (lo < hi) {  // columns ascend within a row
    const int64_t mid = (lo + hi) >> 1;
    const int64_t c = ci[mid];
    if (c == r) { found = mid; break; }
    if (c < r) lo = mid + 1; else hi = mid;
  }
  dpos[r] = found;
}

__global__ void sep_classify(const int64_t* __restrict__ dpos, const double* __restrict__ val,
                             const uint8_t* __restrict__ cons, int64_t n, uint8_t* __restrict__ cls,
                             uint64_t* __restrict__ flag, int64_t* __restrict__ key, int64_t* __restrict__ cnt) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r > n) return;
  if (r == n) {  // totals slot of the scans
    flag[n] = 0;
    cnt[2 * n] = 0;
    return;
  }
  const int64_t p = dpos[r];
  const double diag = p >= 0 ? val[p] : 0.0;
  const bool c_r = cons[r] != 0;
  const uint8_t k = diag == 0.0 ? 0 : (c_r ? 2 : 1);
  cls[r] = k;
  flag[r] = uint64_t(k == 1) | (uint64_t(k == 2) << 32);
  key[r] = (diag == 0.0 && c_r) ? r : INT64_MAX;
  cnt[2 * r] = 0;
  cnt[2 * r + 1] = 0;
}

// Row-count slots: K_aa rows at [0, na), K_ab at [na, 2na), K_ba at
// [2na, 2na + nb), K_bb at [2na + nb, 2na + 2nb).
template <int G>
__global__ void sep_count(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                          const double* __restrict__ val, const uint8_t* __restrict__ cls,
                          const uint64_t* __restrict__ pos, const double* __restrict__ force,
                          const double* __restrict__ disp, int64_t n, int64_t* __restrict__ cnt,
                          int64_t* __restrict__ aa_idx, int64_t* __restrict__ bb_idx, double* __restrict__ r_a,
                          double* __restrict__ u_b) {
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const int lig = threadIdx.x & (G - 1);
  if (r >= n) return;
  const uint8_t k = cls[r];
  if (k == 0) return;
  int ca = 0, cb = 0;
  for (int64_t e = rp[r] + lig; e < rp[r + 1]; e += G) {
    if (val[e] == 0.0) continue;
    const uint8_t kc = cls[ci[e]];
    ca += kc == 1;
    cb += kc == 2;
  }
  ca = group_sum<G>(ca);
  cb = group_sum<G>(cb);
  if (lig != 0) return;
  const int64_t na = lo32(pos[n]), nb = hi32(pos[n]);
  if (k == 1) {
    const int64_t pa = lo32(pos[r]);
    cnt[pa] = ca;
    cnt[na + pa] = cb;
    aa_idx[pa] = r;
    r_a[pa] = force[r];
  } else {
    const int64_t pb = hi32(pos[r]);
    cnt[2 * na + pb] = ca;
    cnt[2 * na + nb + pb] = cb;
    bb_idx[pb] = r;
    u_b[pb] = disp[r];
  }
}

template <int G>
__global__ void sep_fill(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                         const double* __restrict__ val, const uint8_t* __restrict__ cls,
                         const uint64_t* __restrict__ pos, const int64_t* __restrict__ off, int64_t n,
                         int32_t* __restrict__ out_ci, double* __restrict__ out_v) {
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const int lane = threadIdx.x & 63, lig = threadIdx.x & (G - 1);
  if (r >= n) return;
  const uint8_t k = cls[r];
  if (k == 0) return;
  const int64_t na = lo32(pos[n]), nb = hi32(pos[n]);
  int64_t o_a, o_b;  // next slot of an entry in an a column / a b column
  if (k == 1) {
    const int64_t pa = lo32(pos[r]);
    o_a = off[pa];
    o_b = off[na + pa];
  } else {
    const int64_t pb = hi32(pos[r]);
    o_a = off[2 * na + pb];
    o_b = off[2 * na + nb + pb];
  }
  const uint64_t group = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (lane & ~(G - 1));
  const uint64_t below = (1ull << lane) - 1ull;
  const int64_t end = rp[r + 1];
  for (int64_t base = rp[r]; base < end; base += G) {
    const int64_t e = base + lig;
    double v = 0.0;
    int32_t c = 0;
    uint8_t kc = 0;
    if (e < end) {
      v = val[e];
      c = ci[e];
      if (v != 0.0) kc = cls[c];
    }
    const uint64_t ba = __ballot(kc == 1) & group, bb = __ballot(kc == 2) & group;
    if (kc == 1) {
      const int64_t o = o_a + __popcll(ba & below);
      out_ci[o] = int32_t(lo32(pos[c]));
      out_v[o] = v;
    } else if (kc == 2) {
      const int64_t o = o_b + __popcll(bb & below);
      out_ci[o] = int32_t(hi32(pos[c]));
      out_v[o] = v;
    }
    o_a += __popcll(ba);
    o_b += __popcll(bb);
  }
}

// ---- four values per lane (values under 4 GiB, the default)
// One 4-byte word per dof replaces the class byte and the 8-byte scan position
// that the column gathers above read: class << 30 | the dof's index among the
// a dofs (class 1) or the b dofs (class 2); 0 for an excluded dof.
constexpr uint32_t SEP_IDX = (1u << 30) - 1u;

__global__ void sep_pack(const uint8_t* __restrict__ cls, const uint64_t* __restrict__ pos, int64_t n,
                         uint32_t* __restrict__ pack) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t k = cls[r];
  pack[r] = k == 0 ? 0u : (k << 30) | (k == 1 ? lo32(pos[r]) : hi32(pos[r]));
}

// QL lanes per row, each taking four neighbouring entries per step: two
// 16-byte value loads and one 16-byte column load, range-checked buffer loads
// (entries past the row end are loaded but not used; their columns are valid
// indices or 0), and one 4-byte gather per non-zero entry.  On plate and
// beam-frame rows of 18-42 entries the count runs fastest with 4 lanes and
// the fill, whose stores then cover wider spans, with 8 (1024^2 plate: count
// 460 vs 531 us, fill 807 vs 1039 us); 1, 2 and 16 lanes were slower for both.
constexpr int SEP_QL_COUNT = 4;
constexpr int SEP_QL_FILL = 8;
typedef unsigned int sep_u32x4 __attribute__((ext_vector_type(4)));

__device__ inline void sep_load4(__amdgpu_buffer_rsrc_t rv, __amdgpu_buffer_rsrc_t rc, int64_t e, double w[4],
                                 uint32_t c[4]) {
  const uint32_t ev = uint32_t(e);
  const double2 w0 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rv, 8u * ev, 0, 0));
  const double2 w1 = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rv, 8u * ev + 16u, 0, 0));
  const sep_u32x4 cc = __builtin_amdgcn_raw_buffer_load_b128(rc, 4u * ev, 0, 0);
  w[0] = w0.x; w[1] = w0.y; w[2] = w1.x; w[3] = w1.y;
  c[0] = cc.x; c[1] = cc.y; c[2] = cc.z; c[3] = cc.w;
}

template <int SEP_QL>
__global__ void sep_count_quad(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               const double* __restrict__ val, const uint32_t* __restrict__ pack,
                               const uint64_t* __restrict__ pos, const double* __restrict__ force,
                               const double* __restrict__ disp, int64_t n, int64_t* __restrict__ cnt,
                               int64_t* __restrict__ aa_idx, int64_t* __restrict__ bb_idx, double* __restrict__ r_a,
                               double* __restrict__ u_b, uint32_t v_bytes) {
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SEP_QL;
  const int lig = threadIdx.x & (SEP_QL - 1);
  if (r >= n) return;
  const uint32_t pr = pack[r];
  const uint32_t k = pr >> 30;
  if (k == 0) return;
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(val), 0, int(v_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(ci), 0, int(v_bytes / 2u), 0x00020000);
  const int64_t end = rp[r + 1];
  int ca = 0, cb = 0;
  for (int64_t e = rp[r] + 4 * lig; e < end; e += 4 * SEP_QL) {
    double w[4];
    uint32_t c[4];
    sep_load4(rv, rc, e, w, c);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t kc = (e + j < end && w[j] != 0.0) ? pack[c[j]] >> 30 : 0u;
      ca += kc == 1u;
      cb += kc == 2u;
    }
  }
  ca = group_sum<SEP_QL>(ca);
  cb = group_sum<SEP_QL>(cb);
  if (lig != 0) return;
  const int64_t na = lo32(pos[n]), nb = hi32(pos[n]);
  const int64_t pi = int64_t(pr & SEP_IDX);
  if (k == 1) {
    cnt[pi] = ca;
    cnt[na + pi] = cb;
    aa_idx[pi] = r;
    r_a[pi] = force[r];
  } else {
    cnt[2 * na + pi] = ca;
    cnt[2 * na + nb + pi] = cb;
    bb_idx[pi] = r;
    u_b[pi] = disp[r];
  }
}

// Output slots in entry order: an exclusive scan of the per-lane counts over
// the row's SEP_QL lanes (lane l holds entries 4l .. 4l + 3 of the step), then
// each lane's entries in turn.
template <int SEP_QL>
__global__ void sep_fill_quad(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                              const double* __restrict__ val, const uint32_t* __restrict__ pack,
                              const uint64_t* __restrict__ pos, const int64_t* __restrict__ off, int64_t n,
                              int32_t* __restrict__ out_ci, double* __restrict__ out_v, uint32_t v_bytes) {
  const int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SEP_QL;
  const int lig = threadIdx.x & (SEP_QL - 1);
  if (r >= n) return;
  const uint32_t pr = pack[r];
  const uint32_t k = pr >> 30;
  if (k == 0) return;
  const int64_t na = lo32(pos[n]), nb = hi32(pos[n]);
  const int64_t pi = int64_t(pr & SEP_IDX);
  int64_t o_a = k == 1 ? off[pi] : off[2 * na + pi];            // next slot of an entry in an a column
  int64_t o_b = k == 1 ? off[na + pi] : off[2 * na + nb + pi];  // ... in a b column
  const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(val), 0, int(v_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t rc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(ci), 0, int(v_bytes / 2u), 0x00020000);
  // outputs: the four matrices' values and columns together hold at most nnz entries
  const __amdgpu_buffer_rsrc_t ov = __builtin_amdgcn_make_buffer_rsrc(out_v, 0, int(v_bytes), 0x00020000);
  const __amdgpu_buffer_rsrc_t oc = __builtin_amdgcn_make_buffer_rsrc(out_ci, 0, int(v_bytes / 2u), 0x00020000);
  const int64_t end = rp[r + 1];
  for (int64_t base = rp[r]; base < end; base += 4 * SEP_QL) {
    const int64_t e = base + 4 * lig;
    double w[4];
    uint32_t c[4], q[4];
    sep_load4(rv, rc, e, w, c);
    int la = 0, lb = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      q[j] = (e + j < end && w[j] != 0.0) ? pack[c[j]] : 0u;
      la += (q[j] >> 30) == 1u;
      lb += (q[j] >> 30) == 2u;
    }
    int sa = la, sb = lb;  // inclusive scan over the row's lanes
#pragma unroll
    for (int o = 1; o < SEP_QL; o <<= 1) {
      const int ta = __shfl_up(sa, o, SEP_QL), tb = __shfl_up(sb, o, SEP_QL);
      if (lig >= o) {
        sa += ta;
        sb += tb;
      }
    }
    const int tot_a = __shfl(sa, SEP_QL - 1, SEP_QL), tot_b = __shfl(sb, SEP_QL - 1, SEP_QL);
    int64_t ia = o_a + (sa - la), ib = o_b + (sb - lb);
    if (la == 4) {
      // all four entries kept in an a column (nearly every lane: the b dofs
      // are the restrained ones): four consecutive slots, written as two
      // 16-byte value stores and one 16-byte column store, so that a wave's
      // stores cover their span densely
      const uint32_t sa8 = 8u * uint32_t(ia), sa4 = 4u * uint32_t(ia);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(sep_u32x4, make_double2(w[0], w[1])), ov, sa8, 0, 0);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(sep_u32x4, make_double2(w[2], w[3])), ov, sa8 + 16u, 0, 0);
      sep_u32x4 cc;
      cc.x = q[0] & SEP_IDX;
      cc.y = q[1] & SEP_IDX;
      cc.z = q[2] & SEP_IDX;
      cc.w = q[3] & SEP_IDX;
      __builtin_amdgcn_raw_buffer_store_b128(cc, oc, sa4, 0, 0);
      o_a += tot_a;
      o_b += tot_b;
      continue;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t kc = q[j] >> 30;
      if (kc == 1u) {
        out_ci[ia] = int32_t(q[j] & SEP_IDX);
        out_v[ia] = w[j];
        ++ia;
      } else if (kc == 2u) {
        out_ci[ib] = int32_t(q[j] & SEP_IDX);
        out_v[ib] = w[j];
        ++ib;
      }
    }
    o_a += tot_a;
    o_b += tot_b;
  }
}

// Matrix w's row pointers start at out_rp[start_w + w] (rows_w + 1 entries, first 0).
__global__ void sep_finish(const uint64_t* __restrict__ pos, const int64_t* __restrict__ off,
                           const int64_t* __restrict__ err, int64_t n, int64_t* __restrict__ out_rp,
                           int64_t* __restrict__ summary) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t na = lo32(pos[n]), nb = hi32(pos[n]);
  const int64_t start[4] = {0, na, 2 * na, 2 * na + nb};
  const int64_t rows[4] = {na, na, nb, nb};
  if (t == 0) {
    summary[0] = err[0];
    summary[1] = na;
    summary[2] = nb;
    for (int w = 0; w < 4; ++w) {
      summary[3 + w] = off[start[w] + rows[w]] - off[start[w]];
      summary[7 + w] = off[start[w]];
    }
    summary[11] = 0;
  }
  if (t >= 2 * (na + nb) + 4) return;
  int w = 0;
  while (w < 3 && t >= start[w + 1] + w + 1) ++w;
  const int64_t i = t - start[w] - w;  // 0 .. rows_w
  out_rp[t] = off[start[w] + i] - off[start[w]];
}

}  // namespace

std::string upload_bc(Device* d, const uint8_t* constrained, const double* disp, const double* force, size_t n) {
  std::string e;
  if (!(e = dgrow(d->sep_cons, d->sep_cap[0], n)).empty()) return e;
  if (!(e = dgrow(d->sep_disp, d->sep_cap[1], n)).empty()) return e;
  if (!(e = dgrow(d->sep_force, d->sep_cap[2], n)).empty()) return e;
  if (n) {
    HIP_TRY(hipMemcpy(d->sep_cons, constrained, n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->sep_disp, disp, n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d->sep_force, force, n * sizeof(double), hipMemcpyHostToDevice));
  }
  d->sep_bc_rows = int64_t(n);
  return "";
}

std::string separate_sparse(Device* d, SepResult& res, void* stream, bool timed, int lanes_opt) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = d->n_rows;
  res = SepResult();
  d->sep_done = false;
  d->sep_ms = 0.0;
  if (n >= int64_t(INT32_MAX) / 2) return "too many rows for the separation";
  if (lanes_opt != 0 && lanes_opt != 4 && lanes_opt != 8 && lanes_opt != 16 && lanes_opt != 32)
    return "lanes per row must be 0 (automatic), 4, 8, 16 or 32";
  if (d->sep_bc_rows != n) return "boundary conditions are not uploaded";
  if (n == 0) return "";  // no dof: no restraints
  const size_t un = size_t(n);
  std::string e;
  if (!(e = dgrow(d->sep_cls, d->sep_cap[3], un)).empty()) return e;
  if (!(e = dgrow(d->sep_flag, d->sep_cap[4], un + 1)).empty()) return e;
  if (!(e = dgrow(d->sep_pos, d->sep_cap[5], un + 1)).empty()) return e;
  if (!(e = dgrow(d->sep_key, d->sep_cap[6], un)).empty()) return e;
  if (!(e = dgrow(d->sep_cnt, d->sep_cap[7], 2 * un + 1)).empty()) return e;
  if (!(e = dgrow(d->sep_off, d->sep_cap[8], 2 * un + 1)).empty()) return e;
  if (!(e = dgrow(d->sep_aa, d->sep_cap[9], un)).empty()) return e;
  if (!(e = dgrow(d->sep_bb, d->sep_cap[10], un)).empty()) return e;
  if (!(e = dgrow(d->sep_ra, d->sep_cap[11], un)).empty()) return e;
  if (!(e = dgrow(d->sep_ub, d->sep_cap[12], un)).empty()) return e;
  if (!(e = dgrow(d->sep_rp, d->sep_cap[13], 2 * un + 4)).empty()) return e;
  if (!(e = dgrow(d->sep_ci, d->sep_cap[14], size_t(d->nnz))).empty()) return e;
  if (!(e = dgrow(d->sep_v, d->sep_cap[15], size_t(d->nnz))).empty()) return e;
  if (!(e = dgrow(d->sep_sum, d->sep_cap[16], size_t(SEP_SUMMARY + 1))).empty()) return e;
  // scratch of the hipCUB reduction and scans: the largest of the three
  size_t b0 = 0, b1 = 0, b2 = 0;
  HIP_TRY(hipcub::DeviceReduce::Min(nullptr, b0, d->sep_key, d->sep_sum + SEP_SUMMARY, int(n), s));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, d->sep_flag, d->sep_pos, int(n + 1), s));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, d->sep_cnt, d->sep_off, int(2 * n + 1), s));
  if (!(e = dgrow(d->sep_tmp, d->sep_cap[17], std::max(b0, std::max(b1, b2)))).empty()) return e;
  const unsigned B = 256;
  if (!d->sep_dpos_valid) {
    if (!(e = dgrow(d->sep_dpos, d->sep_cap[18], un)).empty()) return e;
    sep_diag_pos<<<unsigned((n + B - 1) / B), B, 0, s>>>(d->row_ptr, d->col_idx, n, d->sep_dpos);
    HIP_TRY(hipGetLastError());
    d->sep_dpos_valid = true;
  }
  if (timed) HIP_TRY(hipEventRecord(d->ev4, s));
  // lanes per row: 8 by default (4 for rows shorter than 8 entries on average);
  // measured faster than 4, 16 or 32 on plate, frame and lattice meshes
  // default: four entries per lane (sep_*_quad) while the values span under
  // 4 GiB; an explicit lanes option (or larger values) takes the one-entry
  // kernels
  const bool quad = lanes_opt == 0 && size_t(d->nnz) * 8 < (size_t(1) << 32);
  const uint32_t v_bytes = quad ? uint32_t(size_t(d->nnz) * 8) : 0u;
  if (quad && !(e = dgrow(d->sep_pack, d->sep_cap[19], un)).empty()) return e;
  const unsigned GQC = unsigned((n * SEP_QL_COUNT + B - 1) / B), GQF = unsigned((n * SEP_QL_FILL + B - 1) / B);
  const int lanes = lanes_opt ? lanes_opt : (d->nnz / n >= 8 ? 8 : 4);
  const unsigned G = unsigned((n * lanes + B - 1) / B);
  auto count = lanes == 32 ? sep_count<32> : lanes == 16 ? sep_count<16> : lanes == 8 ? sep_count<8> : sep_count<4>;
  auto fill = lanes == 32 ? sep_fill<32> : lanes == 16 ? sep_fill<16> : lanes == 8 ? sep_fill<8> : sep_fill<4>;
  sep_classify<<<unsigned((n + 1 + B - 1) / B), B, 0, s>>>(d->sep_dpos, d->val, d->sep_cons, n, d->sep_cls,
                                                           d->sep_flag, d->sep_key, d->sep_cnt);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipcub::DeviceReduce::Min(d->sep_tmp, b0, d->sep_key, d->sep_sum + SEP_SUMMARY, int(n), s));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(d->sep_tmp, b1, d->sep_flag, d->sep_pos, int(n + 1), s));
  if (quad) {
    sep_pack<<<unsigned((n + B - 1) / B), B, 0, s>>>(d->sep_cls, d->sep_pos, n, d->sep_pack);
    sep_count_quad<SEP_QL_COUNT><<<GQC, B, 0, s>>>(d->row_ptr, d->col_idx, d->val, d->sep_pack, d->sep_pos, d->sep_force,
                                    d->sep_disp, n, d->sep_cnt, d->sep_aa, d->sep_bb, d->sep_ra, d->sep_ub, v_bytes);
  } else {
    count<<<G, B, 0, s>>>(d->row_ptr, d->col_idx, d->val, d->sep_cls, d->sep_pos, d->sep_force, d->sep_disp, n,
                          d->sep_cnt, d->sep_aa, d->sep_bb, d->sep_ra, d->sep_ub);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(d->sep_tmp, b2, d->sep_cnt, d->sep_off, int(2 * n + 1), s));
  if (quad)
    sep_fill_quad<SEP_QL_FILL><<<GQF, B, 0, s>>>(d->row_ptr, d->col_idx, d->val, d->sep_pack, d->sep_pos, d->sep_off, n, d->sep_ci,
                                   d->sep_v, v_bytes);
  else
    fill<<<G, B, 0, s>>>(d->row_ptr, d->col_idx, d->val, d->sep_cls, d->sep_pos, d->sep_off, n, d->sep_ci, d->sep_v);
  HIP_TRY(hipGetLastError());
  const unsigned GF = unsigned((2 * n + 4 + B - 1) / B);
  sep_finish<<<GF, B, 0, s>>>(d->sep_pos, d->sep_off, d->sep_sum + SEP_SUMMARY, n, d->sep_rp, d->sep_sum);
  HIP_TRY(hipGetLastError());
  if (timed) HIP_TRY(hipEventRecord(d->ev5, s));
  HIP_TRY(hipStreamSynchronize(s));
  int64_t sum[SEP_SUMMARY];
  HIP_TRY(hipMemcpy(sum, d->sep_sum, sizeof(sum), hipMemcpyDeviceToHost));
  res.err_row = sum[0] == INT64_MAX ? -1 : sum[0];
  res.n_aa = sum[1];
  res.n_bb = sum[2];
  for (int w = 0; w < 4; ++w) {
    res.nnz[w] = sum[3 + w];
    d->sep_base[w] = sum[7 + w];
  }
  d->sep_rows[0] = d->sep_rows[1] = res.n_aa;
  d->sep_rows[2] = d->sep_rows[3] = res.n_bb;
  for (int w = 0; w < 4; ++w) d->sep_nnz[w] = res.nnz[w];
  d->sep_done = true;
  if (timed) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev4, d->ev5));
    d->sep_ms = ms;
  }
  return "";
}

std::string separated_csr(Device* d, int which, const int64_t** rp, const int32_t** ci, const double** v) {
  if (which < 0 || which > 3 || !d->sep_done) return "no separation computed";
  const int64_t start[4] = {0, d->sep_rows[0], 2 * d->sep_rows[0], 2 * d->sep_rows[0] + d->sep_rows[2]};
  if (rp) *rp = d->sep_rp + start[which] + which;
  if (ci) *ci = d->sep_ci + d->sep_base[which];
  if (v) *v = d->sep_v + d->sep_base[which];
  return "";
}

std::string separated_vectors(Device* d, const int64_t** aa, const int64_t** bb, const double** r_a,
                              const double** u_b) {
  if (!d->sep_done) return "no separation computed";
  if (aa) *aa = d->sep_aa;
  if (bb) *bb = d->sep_bb;
  if (r_a) *r_a = d->sep_ra;
  if (u_b) *u_b = d->sep_ub;
  return "";
}

void separation_shape(Device* d, int64_t rows[4], int64_t nnz[4]) {
  for (int w = 0; w < 4; ++w) {
    if (rows) rows[w] = d->sep_done ? d->sep_rows[w] : 0;
    if (nnz) nnz[w] = d->sep_done ? d->sep_nnz[w] : 0;
  }
}

void invalidate_separation(Device* d) { d->sep_done = false; }

double last_separate_ms(Device* d) { return d->sep_ms; }

}  // namespace femdev
