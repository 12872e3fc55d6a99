// plate_fused.hip -- fused assembly of flat Mindlin/MITC4 plate meshes.
//
// One workgroup of 256 threads per patch: up to 64 owned nodes (a compact
// group of nodes in Morton order), the elements touching them (<= 64) and the
// owned rows (owned node a, its r-th element in insertion order; <= 128).
//   step 1: lanes (element, Gauss point) compute the point kinematics of every
//           patch element once (det J, dh/dx, dh/dy, MITC4 multipliers) and
//           the element constants (constitutive scalars, shear vectors) -> LDS
//   step 2: one lane per owned row folds the four points into the row block
//           of node a in that element (52 values, plate.rs:513-724)
//   step 3: row blocks staged in LDS (the same LDS as step 1)
//   step 4: one lane per owned node-pair block sums its contributors in
//           element-insertion order and writes its 13 CSR values (14 on the
//           diagonal) once.  No atomics: results are bitwise reproducible.
// Every element frame is exactly the identity (checked when the maps are
// built), so local and global frames coincide and the drilling entry of an
// owned node is the sum of KROT6 over its elements.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int PF_THREADS = 256;
constexpr int PF_ELEMS = 64;                 // elements touching a patch
constexpr int PF_ROWS = 128;                 // owned (node, element) rows
constexpr int PF_NODES = 64;                 // owned nodes
constexpr int PF_BLOCKS = 2 * PF_THREADS;    // owned node-pair blocks
// LDS strides are odd multiples of 16 bytes so that 16-byte accesses of
// neighbouring lanes fall into different banks.
constexpr int PF_GEO = 18;                   // doubles per (element, point): PlatePoint (17) + pad
constexpr int PF_EDAT = 34;                  // doubles per element: 7 constants, pad, v[12], u[12], pad
constexpr int PF_BS = 14;                    // doubles per staged block (13 used)
constexpr int PF_SLOT = 4 * PF_BS + 2;       // doubles per staged row
constexpr int pf_max(int a, int b) { return a > b ? a : b; }
// LDS: element constants | one region reused for point data, staged rows, staged output
constexpr int PF_REGION = pf_max(pf_max(PF_ROWS * PF_SLOT, PF_ELEMS * 4 * PF_GEO), PF_MAX_OUT);
static_assert(PF_THREADS == 4 * PF_ELEMS, "one lane per (element, point)");
static_assert(PF_ELEMS == PF_MAX_ELEMS && PF_ROWS == PF_MAX_ROWS && PF_NODES == PF_MAX_NODES &&
                  PF_BLOCKS == PF_MAX_BLOCKS,
              "host limits");

// Inputs of one patch held in registers, loaded one patch ahead.  Every load
// is unconditional (indices clamped into the patch, which is never empty) so
// that the loads of the next patch stay in flight while this one computes.
struct PatchRegs {
  uint32_t elem_off, node_off, row_off, blk_off, nel, nown, nrows, nblk, total;
  uint4 c;             // lane's element (t / 4): node indices
  uint32_t eid;        // lane's element id
  double X[4][2];      // lane's element: node coordinates (x, y)
  double4 prm;         // lane's element: E, nu, t, k
  uint32_t row;        // lane's owned row
  int64_t base;        // lane's owned node: CSR position of its row 6a
  uint32_t info, boff, ooff;
  uint32_t blk[2];     // lane's node-pair blocks t and t + 256
};

struct PatchArgs {
  const double* __restrict__ xyz;
  const double* __restrict__ params;
  const uint4* __restrict__ hdr;
  const uint4* __restrict__ pconn;
  const uint32_t* __restrict__ peid;
  const uint32_t* __restrict__ rows;
  const int64_t* __restrict__ obase;
  const uint32_t* __restrict__ oinfo;
  const uint32_t* __restrict__ oboff;
  const uint32_t* __restrict__ oooff;
  const uint32_t* __restrict__ blk;
  int n_patches;
  double* __restrict__ val;
  uint32_t val_bytes;
};

struct PatchLds {
  double2 region[PF_REGION / 2];    // point data | staged rows | staged output
  double edat[PF_ELEMS * PF_EDAT];  // element constants
  int stage_p[PF_ROWS];
  int64_t base[PF_NODES];
  uint32_t info[PF_NODES];
  uint32_t boff[PF_NODES];
  uint32_t ooff[PF_NODES + 1];
};
// the position -> owned node table of step 5 sits in the region above the staged output
static_assert(PF_MAX_OUT * 8 + PF_MAX_OUT * 4 <= PF_REGION * 8, "region too small");

__device__ __forceinline__ void pf_load_header(PatchRegs& P, const PatchArgs& A, int patch) {
  const uint4 h0 = A.hdr[3 * patch], h1 = A.hdr[3 * patch + 1], h2 = A.hdr[3 * patch + 2];
  P.elem_off = h0.x; P.node_off = h0.y; P.row_off = h0.z; P.blk_off = h0.w;
  P.nel = h1.x; P.nown = h1.y; P.nrows = h1.z; P.nblk = h1.w;
  P.total = h2.x;
}
__device__ __forceinline__ void pf_load_maps(PatchRegs& P, const PatchArgs& A, int t) {
  const uint32_t ut = uint32_t(t);
  const uint32_t le = min(ut >> 2, P.nel - 1);
  P.c = A.pconn[P.elem_off + le];
  P.eid = A.peid[P.elem_off + le];
  P.row = A.rows[P.row_off + min(ut >> 1, P.nrows - 1)];  // two lanes per owned row
  const uint32_t o = P.node_off + min(ut, P.nown - 1);
  P.base = A.obase[o];
  P.info = A.oinfo[o];
  P.boff = A.oboff[o];
  P.ooff = A.oooff[o];
  P.blk[0] = A.blk[P.blk_off + min(ut, P.nblk - 1)];
  P.blk[1] = A.blk[P.blk_off + min(ut + PF_THREADS, P.nblk - 1)];
}
__device__ __forceinline__ void pf_load_element(PatchRegs& P, const PatchArgs& A) {
  const uint32_t n[4] = {P.c.x, P.c.y, P.c.z, P.c.w};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double2 v = *reinterpret_cast<const double2*>(A.xyz + 3 * size_t(n[a]));
    P.X[a][0] = v.x; P.X[a][1] = v.y;
  }
  P.prm = *reinterpret_cast<const double4*>(A.params + 4 * size_t(P.eid));
}

// One patch: compute `cur`, prefetch `nxt` (the patch gridDim.x further on),
// whose header was loaded by the previous call.  Vector-memory counters
// retire in issue order, so a wait on a load issued after this patch's
// stores would wait for the stores too: every load whose value is needed in
// the next call is issued before the stores (the header of the patch after
// `nxt` just before them, into `cur`, which is then free).
__device__ __forceinline__ bool pf_process(PatchRegs& cur, PatchRegs& nxt, int patch, const PatchArgs& A,
                                           PatchLds& L) {
  const int t = threadIdx.x;
  const int next = patch + int(gridDim.x);
  const bool has_next = next < A.n_patches;
  double* lds = reinterpret_cast<double*>(L.region);
  double2* geo2 = L.region;  // [elem][point][PF_GEO / 2]
  uint32_t* pos_off = reinterpret_cast<uint32_t*>(lds + PF_MAX_OUT);  // byte offset of each staged value

  __syncthreads();  // the previous patch's LDS reads are complete
  if (t < int(cur.nown)) { L.base[t] = cur.base; L.info[t] = cur.info; L.boff[t] = cur.boff; L.ooff[t] = cur.ooff; }
  if (t == 0) L.ooff[cur.nown] = cur.total;

  // ---- step 1: point kinematics of every patch element
  const int le = t >> 2, ip = t & 3;
  if (le < int(cur.nel)) {
    double xl[6];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int a = m < 2 ? m : 3;
      xl[2 * m] = cur.X[a][0] - cur.X[2][0];
      xl[2 * m + 1] = cur.X[a][1] - cur.X[2][1];
    }
    fem::PlatePoint g;
    fem::plate_point(xl, ip, g);
    double2* go = geo2 + (le * 4 + ip) * (PF_GEO / 2);
    go[0] = make_double2(g.w, g.dx[0]);
    go[1] = make_double2(g.dx[1], g.dx[2]);
    go[2] = make_double2(g.dx[3], g.dy[0]);
    go[3] = make_double2(g.dy[1], g.dy[2]);
    go[4] = make_double2(g.dy[3], g.al[0]);
    go[5] = make_double2(g.al[1], g.al[2]);
    go[6] = make_double2(g.al[3], g.be[0]);
    go[7] = make_double2(g.be[1], g.be[2]);
    go[8] = make_double2(g.be[3], 0.0);
    double* ed = L.edat + le * PF_EDAT;
    if (ip == 0) {
      const fem::PlateMat m = fem::plate_material(cur.prm.x, cur.prm.y, cur.prm.z, cur.prm.w);
      ed[0] = m.cm00; ed[1] = m.cm01; ed[2] = m.cm22; ed[3] = m.cb00; ed[4] = m.cb01; ed[5] = m.cb22; ed[6] = m.cs;
    } else {
      fem::PlateShear sv;
      fem::plate_shear_vectors(xl, sv);
      // lane 1: v of nodes 0-3 (first half), lane 2: v (second half) and u (first), lane 3: u (rest)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int kv = 3 * a + i;
          if ((ip == 1 && kv < 8) || (ip == 2 && kv >= 8)) ed[8 + kv] = sv.v[a][i];
          if ((ip == 2 && kv < 4) || (ip == 3 && kv >= 4)) ed[20 + kv] = sv.u[a][i];
        }
    }
  }
  __syncthreads();
  pf_load_maps(nxt, A, t);

  // ---- step 2: owned rows folded over the four points; two lanes per row,
  //      lane half h takes the column nodes q = 2h, 2h + 1 (12 integrals each)
  const int row = t >> 1, h = t & 1;
  double acc[12];
  int p = 0, re = 0;
  if (row < int(cur.nrows)) {
    re = int(cur.row & 255u);
    p = int(cur.row >> 8);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.0;
#pragma unroll 1
    for (int ip = 0; ip < 4; ++ip) {
      const double2* gi = geo2 + (re * 4 + ip) * (PF_GEO / 2);
      double gv[18];
#pragma unroll
      for (int k = 0; k < 9; ++k) { const double2 v = gi[k]; gv[2 * k] = v.x; gv[2 * k + 1] = v.y; }
      const double w = gv[0];
      const double wdx = w * fem::sel4(gv[1], gv[2], gv[3], gv[4], p);
      const double wdy = w * fem::sel4(gv[5], gv[6], gv[7], gv[8], p);
      const double wal = w * fem::sel4(gv[9], gv[10], gv[11], gv[12], p);
      const double wbe = w * fem::sel4(gv[13], gv[14], gv[15], gv[16], p);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int q = 2 * h + c;
        const double dxq = h ? gv[3 + c] : gv[1 + c], dyq = h ? gv[7 + c] : gv[5 + c];
        const double alq = h ? gv[11 + c] : gv[9 + c], beq = h ? gv[15 + c] : gv[13 + c];
        (void)q;
        acc[6 * c + 0] = __builtin_fma(wdx, dxq, acc[6 * c + 0]);
        acc[6 * c + 1] = __builtin_fma(wdx, dyq, acc[6 * c + 1]);
        acc[6 * c + 2] = __builtin_fma(wdy, dxq, acc[6 * c + 2]);
        acc[6 * c + 3] = __builtin_fma(wdy, dyq, acc[6 * c + 3]);
        acc[6 * c + 4] = __builtin_fma(wal, alq, acc[6 * c + 4]);
        acc[6 * c + 5] = __builtin_fma(wbe, beq, acc[6 * c + 5]);
      }
    }
    if (h == 0) L.stage_p[row] = p;
  }
  __syncthreads();

  // ---- step 3: the row's blocks q = 2h, 2h + 1 from the integrals, staged over the point data
  if (row < int(cur.nrows)) {
    const double* ed = L.edat + re * PF_EDAT;
    fem::PlateMat m;
    m.cm00 = ed[0]; m.cm01 = ed[1]; m.cm22 = ed[2]; m.cb00 = ed[3]; m.cb01 = ed[4]; m.cb22 = ed[5]; m.cs = ed[6];
    double vp[3], up[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { vp[i] = m.cs * ed[8 + 3 * p + i]; up[i] = m.cs * ed[20 + 3 * p + i]; }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int q = 2 * h + c;
      double vq[3], uq[3], o[14];
#pragma unroll
      for (int i = 0; i < 3; ++i) { vq[i] = ed[8 + 3 * q + i]; uq[i] = ed[20 + 3 * q + i]; }
      fem::plate_block_entries(acc[6 * c], acc[6 * c + 1], acc[6 * c + 2], acc[6 * c + 3], acc[6 * c + 4],
                               acc[6 * c + 5], m, vp, up, vq, uq, o);
      o[13] = 0.0;
      double2* dst = L.region + (row * PF_SLOT + PF_BS * q) / 2;
#pragma unroll
      for (int k = 0; k < 7; ++k) dst[k] = make_double2(o[2 * k], o[2 * k + 1]);
    }
  }
  __syncthreads();
  pf_load_element(nxt, A);

  // ---- step 4: one lane per owned node-pair block; contributors summed in
  //      element-insertion order, then placed in CSR order
  double bsum[2][13];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t bi = uint32_t(t + j * PF_THREADS);
    if (bi >= cur.nblk) continue;
    const uint32_t d = cur.blk[j];
    const int n = int((d >> 23) & 63u);
    const uint32_t info = L.info[n];
    const int deg = int(info & 255u), sb = int(info >> 24);
    const bool self = (d >> 29) & 1u;
    const int cnt = self ? deg : int(d & 7u);
    for (int m = 0; m < cnt; ++m) {
      int r, q;
      if (self) { r = m; q = L.stage_p[sb + m]; }
      else { r = int((d >> (3 + 5 * m)) & 7u); q = int((d >> (6 + 5 * m)) & 3u); }
      const double2* src = L.region + ((sb + r) * PF_SLOT + PF_BS * q) / 2;
      double v[14];
#pragma unroll
      for (int h = 0; h < 7; ++h) { const double2 w = src[h]; v[2 * h] = w.x; v[2 * h + 1] = w.y; }
#pragma unroll
      for (int i = 0; i < 13; ++i) bsum[j][i] = (m == 0) ? v[i] : bsum[j][i] + v[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t bi = uint32_t(t + j * PF_THREADS);
    if (bi >= cur.nblk) continue;
    const uint32_t d = cur.blk[j];
    const int n = int((d >> 23) & 63u);
    const uint32_t info = L.info[n];
    const int deg = int(info & 255u), nb = int((info >> 8) & 255u);
    const int k = int(bi - L.boff[n]);
    const int o0 = int(L.ooff[n]);
    const uint32_t g0 = uint32_t(L.base[n]) * 8u;  // 32-bit byte offsets (checked when the maps are built)
    const int pos[13] = {2 * k, 2 * k + 1, 2 * nb + 2 * k, 2 * nb + 2 * k + 1,
                         4 * nb + 3 * k, 4 * nb + 3 * k + 1, 4 * nb + 3 * k + 2,
                         7 * nb + 3 * k, 7 * nb + 3 * k + 1, 7 * nb + 3 * k + 2,
                         10 * nb + 3 * k, 10 * nb + 3 * k + 1, 10 * nb + 3 * k + 2};
#pragma unroll
    for (int i = 0; i < 13; ++i) {
      lds[o0 + pos[i]] = bsum[j][i];
      pos_off[o0 + pos[i]] = g0 + 8u * uint32_t(pos[i]);
    }
    if ((d >> 29) & 1u) {
      // drilling row: KROT6 of every element, in insertion order
      double dr = fem::KROT6;
      for (int r = 1; r < deg; ++r) dr += fem::KROT6;
      lds[o0 + 13 * nb] = dr;
      pos_off[o0 + 13 * nb] = g0 + 8u * uint32_t(13 * nb);
    }
  }
  __syncthreads();

  // ---- step 5: the patch's CSR values, 16 buffer stores per lane (lanes past
  //      the end store to an out-of-range offset, which the hardware drops).
  //      All LDS reads are issued before the first store (no branches), so
  //      their latencies overlap.
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A.val, 0, int(A.val_bytes), 0x00020000);
  const uint32_t total = L.ooff[cur.nown];
  {
    const int after = next + int(gridDim.x);
    pf_load_header(cur, A, after < A.n_patches ? after : patch);
  }
  constexpr int NS = PF_MAX_OUT / PF_THREADS;
  uint32_t off[NS];
  u32x2 v[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const uint32_t pos = uint32_t(t + k * PF_THREADS);
    const uint32_t pc = min(pos, total - 1u);  // total >= 1: a patch owns at least one node
    off[k] = pos_off[pc];
    v[k] = __builtin_bit_cast(u32x2, lds[pc]);
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const uint32_t pos = uint32_t(t + k * PF_THREADS);
    __builtin_amdgcn_raw_buffer_store_b64(v[k], rsrc, pos < total ? off[k] : 0xFFFFFFFFu, 0, 0);
  }
  return has_next;
}

__global__ __launch_bounds__(PF_THREADS, 2) void plate_fused_kernel(PatchArgs A) {
  __shared__ PatchLds L;
  int patch = blockIdx.x;
  if (patch >= A.n_patches) return;
  PatchRegs P0, P1;
  pf_load_header(P0, A, patch);
  pf_load_maps(P0, A, threadIdx.x);
  pf_load_element(P0, A);
  const int second = patch + int(gridDim.x);
  pf_load_header(P1, A, second < A.n_patches ? second : patch);
  // Start the loop with nothing in flight, so that at the loop head every
  // path has the same outstanding vector-memory operations (this patch's
  // loads older than the previous patch's 16 stores) and waits on them
  // leave the stores in flight.
  __builtin_amdgcn_s_waitcnt(0);
  while (true) {
    if (!pf_process(P0, P1, patch, A, L)) break;
    patch += int(gridDim.x);
    if (!pf_process(P1, P0, patch, A, L)) break;
    patch += int(gridDim.x);
  }
}

std::string upload_plate_maps(Device* d, const PlateMaps& pm, int wg_per_cu) {
  d->n_patches = pm.n_patches;
  std::string e;
  if (!(e = dupload(d->pf_hdr, pm.hdr.data(), pm.hdr.size())).empty()) return e;
  if (!(e = dupload(d->pf_conn, pm.pconn.data(), pm.pconn.size())).empty()) return e;
  if (!(e = dupload(d->pf_eid, pm.peid.data(), pm.peid.size())).empty()) return e;
  if (!(e = dupload(d->pf_rows, pm.rows.data(), pm.rows.size())).empty()) return e;
  if (!(e = dupload(d->pf_obase, pm.own_base.data(), pm.own_base.size())).empty()) return e;
  if (!(e = dupload(d->pf_oinfo, pm.own_info.data(), pm.own_info.size())).empty()) return e;
  if (!(e = dupload(d->pf_oboff, pm.own_boff.data(), pm.own_boff.size())).empty()) return e;
  if (!(e = dupload(d->pf_oooff, pm.own_ooff.data(), pm.own_ooff.size())).empty()) return e;
  if (!(e = dupload(d->pf_blk, pm.blk.data(), pm.blk.size())).empty()) return e;
  return set_plate_grid(d, wg_per_cu);
}

std::string set_plate_grid(Device* d, int wg_per_cu) {
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->dev));
  d->pf_grid = wg_per_cu <= 0 ? d->n_patches : std::max(1, std::min(d->n_patches, cus * wg_per_cu));
  return "";
}

std::string assemble_plate_fused(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->n_patches > 0) {
    PatchArgs a;
    a.xyz = d->xyz;
    a.params = d->params[PLATE];
    a.hdr = reinterpret_cast<const uint4*>(d->pf_hdr);
    a.pconn = reinterpret_cast<const uint4*>(d->pf_conn);
    a.peid = d->pf_eid;
    a.rows = d->pf_rows;
    a.obase = d->pf_obase;
    a.oinfo = d->pf_oinfo;
    a.oboff = d->pf_oboff;
    a.oooff = d->pf_oooff;
    a.blk = d->pf_blk;
    a.n_patches = d->n_patches;
    a.val = val;
    a.val_bytes = uint32_t(d->nnz * 8);
    plate_fused_kernel<<<d->pf_grid, PF_THREADS, 0, s>>>(a);
    HIP_TRY(hipGetLastError());
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

}  // namespace femdev
