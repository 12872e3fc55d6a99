// bar_owner.hip -- one-pass assembly of truss and beam meshes.
//
// The host cuts the nodes into ranges of consecutive nodes.  Ranges whose
// nodes touch trusses only go to truss_owner_kernel (below: a compact entry
// record, rows 0..2 only); the others to bar_owner_kernel:
//
// One workgroup per range of owned nodes (<= 40 nodes, <= 256 incident
// (node, element) entries, <= 5120 CSR values).  Each entry's constants (the
// element frame and term scales, its blocks' masks and row offsets) sit in a
// 160-byte slot formed once per model change (bar_slot_kernel), in the order
// the workgroup reads them.  One lane per (owned node a, CSR row i) walks a's
// entries in element-insertion order and computes row i of each entry's global
// blocks (p, p) and (p, 1 - p) (p = a's position in the element).  The six
// lanes of a node sit in one wave and load each slot together (two 16-byte
// loads per lane), one entry ahead of the arithmetic, into the node's LDS slot,
// from which every lane reads it.  The diagonal block (a, a) is summed in
// registers over all incident elements (insertion order, starting from the
// first contribution); an off-diagonal block (a, b) has exactly one
// contributing element (checked when the maps are built).  The rows are formed
// in LDS and the workgroup's owned rows, one contiguous CSR range, are then
// written out by consecutive lanes as non-temporal 16-byte stores (a scattered
// row-by-row write costs one memory request per value).  Every CSR value is
// written once: no contribution buffer, no atomics, bitwise reproducible.
//
// Beam blocks use the rank-one form of beam.rs:495-605: with T = diag(R, R),
//   row i of T^T (sum_t s_t v_t,p v_t,q^T) T = sum_t s_t (T^T v_t,p)_i (T^T v_t,q)^T.
#include <hip/hip_runtime.h>

#include <string>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

namespace {

typedef double bo_d2 __attribute__((ext_vector_type(2)));

constexpr int BO_THREADS = 256;
constexpr int BO_NODES = 40;     // owned nodes per workgroup (10 per wave, 6 row lanes each)
constexpr int BO_ENTRIES = 256;  // (owned node, incident element) entries per workgroup
constexpr int BO_DATA = 16;      // element constants per entry slot
constexpr int BO_SLOT = 20;      // doubles per entry slot
constexpr int BO_VALUES = 5120;  // CSR values of the owned rows per workgroup
constexpr int TO_NODES = 85;     // truss kernel: owned nodes per workgroup (3 row lanes each)
constexpr int TO_ENTRIES = 512;  // truss kernel: entries per workgroup

struct BarArgs {
  const double* __restrict__ xyz;
  const uint32_t* __restrict__ conn[2];
  const double* __restrict__ params[2];
  const double* __restrict__ R[2];
  const double* __restrict__ tdir;  // trusses: first rows of R, 3 doubles each
  const double* __restrict__ derived;
  const uint32_t* __restrict__ bid[2];
  const uint64_t* __restrict__ bmask;
  const uint16_t* __restrict__ boff;
  const uint32_t* __restrict__ inc_ptr;
  const uint32_t* __restrict__ inc;
  const uint2* __restrict__ ranges;  // per workgroup: first node, node count
  const double* __restrict__ slot;   // bar kernel: staged entries (BO_SLOT doubles each), range r from slot_ptr[r]
  const uint32_t* __restrict__ slot_ptr;
  const int64_t* __restrict__ row_ptr;
  double* __restrict__ val;
};

struct BarLds {
  uint32_t first[BO_NODES + 1];         // entry range of each owned node
  double2 slot[BO_NODES][BO_SLOT / 2];  // each owned node's current entry slot
  double out[BO_VALUES];                // the owned rows
};
static_assert(3 * sizeof(BarLds) <= 160 * 1024, "three workgroups per CU");

static_assert(BO_NODES == BAR_WG_NODES && BO_ENTRIES == BAR_WG_ENTRIES && BO_VALUES == BAR_WG_VALUES, "range cut");

// Truss-only ranges: rows 3..5 are empty and an entry needs five constants,
// so the layout is compact (four workgroups per CU) and a range holds twice
// the entries.
struct TrussLds {
  double data[TO_ENTRIES][5];      // r (3), kpp, kpq
  uint64_t mask[TO_ENTRIES][2];    // masks of the entry's blocks (p, p) and (p, 1 - p)
  uint16_t off[TO_ENTRIES][2][3];  // offsets of rows 0..2 of those blocks
  uint8_t node[TO_ENTRIES];        // owned node of the entry
  uint32_t first[TO_NODES + 1];
  uint32_t row[TO_NODES][3];       // start of rows 0..2 of each owned node, from the range's first value
};
static_assert(4 * sizeof(TrussLds) <= 160 * 1024, "four workgroups per CU");
static_assert(TO_NODES == TRUSS_WG_NODES && TO_ENTRIES == TRUSS_WG_ENTRIES && 3 * TO_NODES <= BO_THREADS, "range cut");

// T^T v_t for the beam's rank-one term t with nodal shape value b
// (v: u = b e0; v = b e1 - e5/2; w = b e2 - e4/2; tu = b e3; tv = b e4; tw = b e5)
__device__ inline void beam_term_vec(int t, double b, const double R[9], double o[6]) {
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double tr = 0.0, rr = 0.0;
    if (t == 0) tr = b * R[0 * 3 + m];
    else if (t == 1) { tr = b * R[1 * 3 + m]; rr = -0.5 * R[2 * 3 + m]; }
    else if (t == 2) { tr = b * R[2 * 3 + m]; rr = -0.5 * R[1 * 3 + m]; }
    else rr = b * R[(t - 3) * 3 + m];
    o[m] = tr;
    o[3 + m] = rr;
  }
}

// Jacobian of element e of `type` and the nodal shape derivatives b_p, b_q
// of its end p and the other end
__device__ inline double bar_shape(const BarArgs& A, int type, size_t e, int p, double& bp, double& bq) {
  const uint32_t n0 = A.conn[type][2 * e], n1 = A.conn[type][2 * e + 1];
  double d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = A.xyz[3 * size_t(n1) + k] - A.xyz[3 * size_t(n0) + k];
  const double Lg = fem::norm3(d[0], d[1], d[2]);
  const double J = fem::bar_jacobian(Lg);
  const double inv_j = 1.0 / J;
  bp = (p == 0 ? 0.0 - 0.5 : 0.0 + 0.5) * inv_j;
  bq = (p == 0 ? 0.0 + 0.5 : 0.0 - 0.5) * inv_j;
  return J;
}

// truss constants (truss.rs:132-189): direction r and the scales of the
// (p, p) and (p, q) blocks, k_pq = kpq r r^T
__device__ inline void truss_constants(const BarArgs& A, size_t e, int p, double o[5]) {
  double bp, bq;
  const double J = bar_shape(A, 0, e, p, bp, bq);
  const double* Rp = A.tdir + 3 * e;  // R's first row, stored compactly by the setup
  const double* prm = A.params[0] + 3 * e;
  const double E = prm[0], Ar = prm[1], A2 = prm[2];
  const double area = (A2 == A2) ? (A2 - Ar) / 2.0 * 0.0 + Ar - (A2 - Ar) / 2.0 * -1.0 : Ar;
  const double scale = area * E * J * 2.0;
  o[0] = Rp[0]; o[1] = Rp[1]; o[2] = Rp[2];
  o[3] = (bp * bp) * scale;
  o[4] = (bp * bq) * scale;
}

// One (owned node, incident element) entry's constants, as a slot of BO_SLOT
// doubles: the element constants (BO_DATA doubles; beam: R (9), the six term
// scales, b_p = -b_q; truss: r (3), k_pp, k_pq; truss.rs:132-189,
// beam.rs:495-605), then the masks of its blocks (p, p) and (p, 1 - p) (36
// bits each) with the offsets of rows 0 and 1 in the bits above them (12 bits
// each: a row holds at most 6 x 257 values) and the element type in bit 63,
// then the offsets of rows 2..5 of both blocks (16 bits each).  The slots are
// formed once per model change by bar_slot_kernel, so that the assembly reads
// each entry with ten 16-byte loads instead of a chain of dependent ones.
static_assert(BO_DATA + 4 == BO_SLOT && BO_SLOT == 20, "slot layout: ten 16-byte chunks");

__device__ inline void form_slot(const BarArgs& A, uint32_t u, double* __restrict__ slot) {
  const int type = int(u >> 31), p = int(u & 1u);
  const size_t e = (u >> 1) & 0x3FFFFFFFu;
  double o[BO_DATA + 1];
#pragma unroll
  for (int k = 0; k <= BO_DATA; ++k) o[k] = 0.0;
  if (type == 0) {
    truss_constants(A, e, p, o);
  } else {
    double bp, bq;
    const double J = bar_shape(A, 1, e, p, bp, bq);
    const double* Rp = A.R[1] + 9 * e;
    const double* prm = A.params[1] + 11 * e;
    const double E = prm[0], nu = prm[1], Ar = prm[2], it = prm[6], k = prm[7];
    const double G = E / (2.0 * (1.0 + nu));
    const double det_j = J, alpha = 2.0;
    const double i11p = A.derived[2 * e], i22p = A.derived[2 * e + 1];
#pragma unroll
    for (int m = 0; m < 9; ++m) o[m] = Rp[m];
    o[9] = Ar * E * det_j * alpha;
    o[10] = G * Ar * k * det_j * alpha;
    o[11] = G * Ar * k * det_j * alpha;
    o[12] = G * it * det_j * alpha;
    o[13] = E * i22p * det_j * alpha;
    o[14] = E * i11p * det_j * alpha;
    o[15] = bp;
    (void)bq;  // -bp exactly
  }
  const uint32_t ks = A.bid[type][4 * e + 3 * p], kn = A.bid[type][4 * e + 2 * p + (1 - p)];
  uint16_t off[12];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    off[i] = A.boff[6 * size_t(ks) + i];
    off[6 + i] = A.boff[6 * size_t(kn) + i];
  }
  double w[BO_SLOT];
#pragma unroll
  for (int k = 0; k < BO_DATA; ++k) w[k] = o[k];
  // masks (36 bits) with the offsets of rows 0, 1 (12 bits each) above them and
  // the type in bit 63; the offsets of rows 2..5 as 16-bit fields
  const uint64_t m0 = A.bmask[ks], m1 = A.bmask[kn];
  w[BO_DATA] = __builtin_bit_cast(double, m0 | uint64_t(off[0]) << 36 | uint64_t(off[1]) << 48 | uint64_t(type) << 63);
  w[BO_DATA + 1] = __builtin_bit_cast(double, m1 | uint64_t(off[6]) << 36 | uint64_t(off[7]) << 48);
#pragma unroll
  for (int b = 0; b < 2; ++b)
    w[BO_DATA + 2 + b] = __builtin_bit_cast(double, uint64_t(off[6 * b + 2]) | uint64_t(off[6 * b + 3]) << 16 |
                                                        uint64_t(off[6 * b + 4]) << 32 | uint64_t(off[6 * b + 5]) << 48);
  double2* dst = reinterpret_cast<double2*>(slot);
#pragma unroll
  for (int k = 0; k < BO_SLOT / 2; ++k) dst[k] = make_double2(w[2 * k], w[2 * k + 1]);
}

// slot s <- entry slot_inc[s]
__global__ __launch_bounds__(256) void bar_slot_kernel(BarArgs A, const uint32_t* __restrict__ slot_inc, size_t n,
                                                       double* __restrict__ slot) {
  const size_t s = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (s >= n) return;
  form_slot(A, A.inc[slot_inc[s]], slot + BO_SLOT * s);
}

// slot fields: element type, mask of block b (0: (p, p), 1: (p, 1 - p)) and
// the offset of its row i within the owned node's row
__device__ inline uint32_t slot_type(const double* o) {
  return uint32_t(__builtin_bit_cast(uint64_t, o[BO_DATA]) >> 63);
}
__device__ inline uint64_t slot_mask(const double* o, int b) {
  return __builtin_bit_cast(uint64_t, b ? o[BO_DATA + 1] : o[BO_DATA]) & ((uint64_t(1) << 36) - 1);
}
__device__ inline int slot_off(const double* o, int b, int i) {
  const uint64_t lo = __builtin_bit_cast(uint64_t, b ? o[BO_DATA + 1] : o[BO_DATA]);
  const uint64_t hi = __builtin_bit_cast(uint64_t, b ? o[BO_DATA + 3] : o[BO_DATA + 2]);
  return i < 2 ? int((lo >> (36 + 12 * i)) & 0xFFFu) : int((hi >> (16 * (i - 2))) & 0xFFFFu);
}

// row i of an entry's blocks (p, p) -> self and (p, 1 - p) -> nb
__device__ inline void entry_rows(const double* __restrict__ o, uint32_t type, int i, double self[6], double nb[6]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) { self[j] = 0.0; nb[j] = 0.0; }
  if (type == 0) {
    if (i >= 3) return;
    const double ri = i == 0 ? o[0] : (i == 1 ? o[1] : o[2]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      self[j] = (ri * o[3]) * o[j];
      nb[j] = (ri * o[4]) * o[j];
    }
    return;
  }
  double R[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) R[m] = o[m];
  const double bp = o[15], bq = -bp;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    double vp[6], vq[6];
    beam_term_vec(t, bp, R, vp);
    beam_term_vec(t, bq, R, vq);
    const double ai = i == 0 ? vp[0] : (i == 1 ? vp[1] : (i == 2 ? vp[2] : (i == 3 ? vp[3] : (i == 4 ? vp[4] : vp[5]))));
    const double sa = o[9 + t] * ai;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      self[j] += sa * vp[j];
      nb[j] += sa * vq[j];
    }
  }
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row i of one block into the staged rows (r0: the row's start within them)
__device__ inline void write_row(BarLds& L, uint64_t mask, int off_i, int i, int r0, const double v[6]) {
  const uint32_t mi = uint32_t(mask >> (6 * i)) & 63u;
  const int d0 = r0 + off_i;
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if ((mi >> j) & 1u) L.out[d0 + __builtin_popcount(mi & ((1u << j) - 1u))] = v[j];
}

// One workgroup per range of owned nodes: one lane per (owned node, row) forms
// the rows in LDS, each entry's slot loaded during the previous entry's
// arithmetic; then consecutive lanes write the range out.
__global__ __launch_bounds__(BO_THREADS) void bar_owner_kernel(BarArgs A) {
  __shared__ BarLds L;
  const int t = threadIdx.x;
  const uint2 rg = A.ranges[blockIdx.x];
  const uint32_t a0 = rg.x, na = rg.y;
  const uint32_t e0 = A.inc_ptr[a0];
  const int64_t base = A.row_ptr[6 * size_t(a0)];
  const int nv = int(A.row_ptr[6 * size_t(a0 + na)] - base);  // <= BO_VALUES (cut on the host)
  if (t <= int(na)) L.first[t] = A.inc_ptr[a0 + t] - e0;
  __syncthreads();
  const double2* __restrict__ slots =
      reinterpret_cast<const double2*>(A.slot + BO_SLOT * size_t(A.slot_ptr[blockIdx.x]));
  // lane (owned node j, row i), ten nodes per wave (lanes 60..63 idle here).
  // The six lanes of a node load one entry slot together (lane i: 16-byte
  // chunks i and i + 6) into the node's LDS slot, one entry ahead of the
  // arithmetic, and each reads the whole slot from there.
  const int jl = (t & 63) / 6, i = (t & 63) - 6 * jl;
  const int j = 10 * (t >> 6) + jl;
  if (jl < 10 && j < int(na)) {
    const size_t a = size_t(a0) + j;
    const int r0 = int(A.row_ptr[6 * a + i] - base);
    const bool row = A.row_ptr[6 * a + i + 1] - base > r0;
    const int s0 = int(L.first[j]), s1 = int(L.first[j + 1]);
    double2* __restrict__ buf = L.slot[j];
    double2 c0 = make_double2(0.0, 0.0), c1 = c0;
    if (s1 > s0) {
      c0 = slots[(BO_SLOT / 2) * s0 + i];
      if (i < 4) c1 = slots[(BO_SLOT / 2) * s0 + 6 + i];
      buf[i] = c0;
      if (i < 4) buf[6 + i] = c1;
      if (s0 + 1 < s1) {
        c0 = slots[(BO_SLOT / 2) * (s0 + 1) + i];
        if (i < 4) c1 = slots[(BO_SLOT / 2) * (s0 + 1) + 6 + i];
      }
    }
    wave_lds_sync();
    double acc[6];
    uint64_t m0 = 0;  // block (a, a): mask and row offset, from the first entry
    int o0 = 0;
    for (int sl = s0; sl < s1; ++sl) {
      double o[BO_SLOT];
#pragma unroll
      for (int m = 0; m < BO_SLOT / 2; ++m) {
        const double2 v = buf[m];
        o[2 * m] = v.x;
        o[2 * m + 1] = v.y;
      }
      if (row) {
        double self[6], nb[6];
        entry_rows(o, slot_type(o), i, self, nb);
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] = sl == s0 ? self[q] : acc[q] + self[q];
        if (sl == s0) {
          m0 = slot_mask(o, 0);
          o0 = slot_off(o, 0, i);
        }
        write_row(L, slot_mask(o, 1), slot_off(o, 1, i), i, r0, nb);
      }
      if (sl + 1 < s1) {
        wave_lds_sync();  // every lane of the node has read the slot
        buf[i] = c0;
        if (i < 4) buf[6 + i] = c1;
        wave_lds_sync();
        if (sl + 2 < s1) {
          c0 = slots[(BO_SLOT / 2) * (sl + 2) + i];
          if (i < 4) c1 = slots[(BO_SLOT / 2) * (sl + 2) + 6 + i];
        }
      }
    }
    if (row && s1 > s0) write_row(L, m0, o0, i, r0, acc);
  }
  __syncthreads();
  double* __restrict__ dst = A.val + base;
  // 16-byte stores from the first even CSR position, non-temporal (nothing in
  // this kernel reads the values again); the odd values at either end alone
  const int head = int(base & 1), npair = (nv - head) >> 1;
  for (int k = t; k < npair; k += BO_THREADS) {
    const int o = head + 2 * k;
    bo_d2 v;
    v.x = L.out[o];
    v.y = L.out[o + 1];
    __builtin_nontemporal_store(v, reinterpret_cast<bo_d2*>(dst + o));
  }
  if (t == 0 && head && nv > 0) dst[0] = L.out[0];
  if (t == BO_THREADS - 1 && nv > head && ((nv - head) & 1)) dst[nv - 1] = L.out[nv - 1];
}

// row i < 3 of a truss block (columns 0..2) straight to the CSR
__device__ inline void write_row3(double* val, uint64_t mask, int off, int i, int64_t r0, const double v[3]) {
  const uint32_t mi = uint32_t(mask >> (6 * i)) & 7u;
  double* dst = val + (r0 + off);
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if ((mi >> j) & 1u) dst[__builtin_popcount(mi & ((1u << j) - 1u))] = v[j];
}

// one truss entry's constants, masks and row offsets (registers)
struct TrussEntry {
  double o[5];
  uint64_t m[2];
  uint16_t off[2][3];
};

__device__ inline void truss_entry(const BarArgs& A, uint32_t u, TrussEntry& E) {
  const int p = int(u & 1u);
  const size_t e = (u >> 1) & 0x3FFFFFFFu;
  truss_constants(A, e, p, E.o);
  const uint32_t ks = A.bid[0][4 * e + 3 * p], kn = A.bid[0][4 * e + 2 * p + (1 - p)];
  E.m[0] = A.bmask[ks];
  E.m[1] = A.bmask[kn];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    E.off[0][i] = A.boff[6 * size_t(ks) + i];
    E.off[1][i] = A.boff[6 * size_t(kn) + i];
  }
}

__device__ inline void put_entry(TrussLds& L, int sl, const TrussEntry& E) {
#pragma unroll
  for (int k = 0; k < 5; ++k) L.data[sl][k] = E.o[k];
  L.mask[sl][0] = E.m[0];
  L.mask[sl][1] = E.m[1];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    L.off[sl][0][i] = E.off[0][i];
    L.off[sl][1][i] = E.off[1][i];
  }
}

// Truss-only ranges: the same sums as bar_owner_kernel (diagonal block summed
// in insertion order from the first contribution, off-diagonal blocks written
// by their one element).  Staging: two entries per lane, loads of both in
// flight together.  Off-diagonal rows: one task per (row, entry), so that
// consecutive lanes write neighbouring blocks of one row.  Diagonal rows: one
// lane per (owned node, row 0..2).
__global__ __launch_bounds__(BO_THREADS) void truss_owner_kernel(BarArgs A) {
  __shared__ TrussLds L;
  const int t = threadIdx.x;
  const uint2 rg = A.ranges[blockIdx.x];
  const uint32_t a0 = rg.x, na = rg.y;
  const uint32_t e0 = A.inc_ptr[a0];
  const int ne = int(A.inc_ptr[a0 + na] - e0);  // <= TO_ENTRIES (cut on the host)
  const int64_t base = A.row_ptr[6 * size_t(a0)];
  if (t <= int(na)) L.first[t] = A.inc_ptr[a0 + t] - e0;
  if (t < int(na)) {
    const int f0 = int(A.inc_ptr[a0 + t] - e0), f1 = int(A.inc_ptr[a0 + t + 1] - e0);
    for (int sl = f0; sl < f1; ++sl) L.node[sl] = uint8_t(t);
  }
  if (t < 3 * int(na)) L.row[t / 3][t % 3] = uint32_t(A.row_ptr[6 * size_t(a0 + t / 3) + t % 3] - base);
  if (ne > 0) {
    TrussEntry E0, E1;
    truss_entry(A, A.inc[e0 + min(t, ne - 1)], E0);
    truss_entry(A, A.inc[e0 + min(t + BO_THREADS, ne - 1)], E1);
    if (t < ne) put_entry(L, t, E0);
    if (t + BO_THREADS < ne) put_entry(L, t + BO_THREADS, E1);
  }
  __syncthreads();
  double* __restrict__ val = A.val + base;
  for (int k = t; k < 3 * ne; k += BO_THREADS) {
    const int i = k / ne, sl = k - i * ne;
    const double* o = L.data[sl];
    const double s = o[i] * o[4];
    double nb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) nb[q] = s * o[q];
    write_row3(val, L.mask[sl][1], L.off[sl][1][i], i, L.row[L.node[sl]][i], nb);
  }
  const int j = t / 3, i = t % 3;
  if (j >= int(na)) return;
  const int s0 = int(L.first[j]), s1 = int(L.first[j + 1]);
  if (s1 == s0) return;
  double acc[3];
  for (int sl = s0; sl < s1; ++sl) {
    const double* o = L.data[sl];
    const double s = o[i] * o[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = sl == s0 ? s * o[q] : acc[q] + s * o[q];
  }
  write_row3(val, L.mask[s0][0], L.off[s0][0][i], i, L.row[j][i], acc);
}

}  // namespace

namespace {

BarArgs bar_args(Device* d, double* val) {
  BarArgs A;
  A.xyz = d->xyz;
  for (int t = 0; t < 2; ++t) {
    A.conn[t] = d->conn[t];
    A.params[t] = d->params[t];
    A.R[t] = d->R[t];
    A.bid[t] = d->bo_bid[t];
  }
  A.derived = d->derived;
  A.tdir = d->truss_dir;
  A.bmask = d->bo_bmask;
  A.boff = d->bo_boff;
  A.inc_ptr = d->bo_inc_ptr;
  A.inc = d->bo_inc;
  A.ranges = reinterpret_cast<const uint2*>(d->bo_ranges);
  A.slot = d->bo_slot;
  A.slot_ptr = d->bo_slot_ptr;
  A.row_ptr = d->row_ptr;
  A.val = val;
  return A;
}

// (re)forms the bar kernel's entry slots from the current coordinates,
// frames and parameters
std::string fill_bar_slots(Device* d, hipStream_t s) {
  if (d->bo_n_slots) {
    BarArgs A = bar_args(d, nullptr);
    bar_slot_kernel<<<unsigned((d->bo_n_slots + 255) / 256), 256, 0, s>>>(A, d->bo_slot_inc, d->bo_n_slots,
                                                                          d->bo_slot);
    HIP_TRY(hipGetLastError());
  }
  d->bo_slot_dirty = false;
  return "";
}

}  // namespace

std::string upload_bar_maps(Device* d, const BarMaps& bm) {
  std::string e;
  for (int t = 0; t < 2; ++t)
    if (!(e = dupload(d->bo_bid[t], bm.bid[t].data(), bm.bid[t].size())).empty()) return e;
  if (!(e = dupload(d->bo_bmask, bm.bmask.data(), bm.bmask.size())).empty()) return e;
  if (!(e = dupload(d->bo_boff, bm.boff.data(), bm.boff.size())).empty()) return e;
  if (!(e = dupload(d->bo_inc_ptr, bm.inc_ptr.data(), bm.inc_ptr.size())).empty()) return e;
  if (!(e = dupload(d->bo_ranges, bm.ranges.data(), bm.ranges.size())).empty()) return e;
  d->bo_n_ranges = bm.n_bar_ranges;
  d->bo_n_truss_ranges = bm.ranges.size() / 2 - bm.n_bar_ranges;
  if (!(e = dupload(d->bo_inc, bm.inc.data(), bm.inc.size())).empty()) return e;
  if (!(e = dupload(d->bo_slot_inc, bm.slot_inc.data(), bm.slot_inc.size())).empty()) return e;
  if (!(e = dupload(d->bo_slot_ptr, bm.slot_ptr.data(), bm.slot_ptr.size())).empty()) return e;
  d->bo_n_slots = bm.slot_inc.size();
  if (!(e = dalloc(d->bo_slot, size_t(BO_SLOT) * d->bo_n_slots)).empty()) return e;
  e = fill_bar_slots(d, nullptr);
  if (!e.empty()) return e;
  HIP_TRY(hipDeviceSynchronize());
  return "";
}

std::string assemble_bar_owner(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (d->bo_slot_dirty) {  // coordinates or parameters changed since the slots were formed
    auto e = fill_bar_slots(d, s);
    if (!e.empty()) return e;
  }
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->bo_n_ranges + d->bo_n_truss_ranges) {
    BarArgs A = bar_args(d, val);
    if (d->bo_n_ranges) {
      bar_owner_kernel<<<unsigned(d->bo_n_ranges), BO_THREADS, 0, s>>>(A);
      HIP_TRY(hipGetLastError());
    }
    if (d->bo_n_truss_ranges) {
      A.ranges += d->bo_n_ranges;
      truss_owner_kernel<<<unsigned(d->bo_n_truss_ranges), BO_THREADS, 0, s>>>(A);
      HIP_TRY(hipGetLastError());
    }
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

}  // namespace femdev
