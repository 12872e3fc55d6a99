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

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

constexpr int PF_THREADS = 256;
constexpr int PF_ELEMS = 64;                 // elements touching a patch
constexpr int PF_ROWS = 128;                 // owned (node, element) rows
constexpr int PF_NODES = 64;                 // owned nodes
constexpr int PF_OUT = PF_MAX_OUT;           // staged CSR values of a patch (13 nb + 1 per node)
// LDS strides are odd multiples of 16 bytes so that 16-byte accesses of
// neighbouring lanes fall into different banks.
constexpr int PF_GEO = 18;                   // doubles per (element, point): PlatePoint (17) + pad
constexpr int PF_EDAT = 34;                  // doubles per element: 7 constants, pad, v[12], u[12], pad
constexpr int PF_BS = 14;                    // doubles per staged block (13 used)
constexpr int PF_SLOT = 4 * PF_BS + 2;       // doubles per staged row
constexpr int PF_STEP1 = PF_ELEMS * (4 * PF_GEO + PF_EDAT);
constexpr int pf_max(int a, int b) { return a > b ? a : b; }
constexpr int PF_LDS = pf_max(pf_max(PF_ROWS * PF_SLOT, PF_STEP1), PF_OUT);
static_assert(PF_THREADS == 4 * PF_ELEMS, "one lane per (element, point)");
static_assert(PF_ELEMS == PF_MAX_ELEMS && PF_ROWS == PF_MAX_ROWS && PF_NODES == PF_MAX_NODES, "host limits");

__global__ __launch_bounds__(PF_THREADS) void plate_fused_kernel(
    const double* __restrict__ xyz, const uint32_t* __restrict__ conn, const double* __restrict__ params,
    const uint32_t* __restrict__ node_ptr, const uint32_t* __restrict__ own_node,
    const uint32_t* __restrict__ own_info, const uint32_t* __restrict__ blk_ptr, const uint32_t* __restrict__ blk,
    const uint32_t* __restrict__ elem_ptr, const uint32_t* __restrict__ elems, const uint32_t* __restrict__ row_ptr_,
    const uint32_t* __restrict__ rows, const int64_t* __restrict__ csr_row_ptr, double* __restrict__ val) {
  __shared__ double2 lds2[PF_LDS / 2];
  __shared__ int stage_p[PF_ROWS];
  __shared__ int64_t s_base[PF_NODES];
  __shared__ uint32_t s_info[PF_NODES];
  __shared__ uint32_t s_boff[PF_NODES];
  __shared__ uint32_t s_ooff[PF_NODES + 1];
  double* lds = reinterpret_cast<double*>(lds2);
  double2* geo2 = lds2;                                   // [elem][point][PF_GEO / 2]
  double* edat = lds + PF_ELEMS * 4 * PF_GEO;             // [elem][PF_EDAT]
  const int patch = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t n0 = node_ptr[patch], nown = node_ptr[patch + 1] - n0;
  const uint32_t e0 = elem_ptr[patch], nel = elem_ptr[patch + 1] - e0;
  const uint32_t r0 = row_ptr_[patch], nrows = row_ptr_[patch + 1] - r0;
  const uint32_t b0 = blk_ptr[n0], nblk = blk_ptr[n0 + nown] - b0;

  if (t < int(nown)) {
    const uint32_t a = own_node[n0 + t];
    s_base[t] = csr_row_ptr[6 * int64_t(a)];
    s_info[t] = own_info[n0 + t];
    s_boff[t] = blk_ptr[n0 + t] - b0;
  }

  // ---- step 1: point kinematics of every patch element
  const int le = t >> 2, ip = t & 3;
  if (le < int(nel)) {
    const uint32_t e = elems[e0 + le];
    const uint4 c = *reinterpret_cast<const uint4*>(conn + 4 * size_t(e));
    const double x3 = xyz[3 * size_t(c.z)], y3 = xyz[3 * size_t(c.z) + 1];
    double xl[6];
    xl[0] = xyz[3 * size_t(c.x)] - x3; xl[1] = xyz[3 * size_t(c.x) + 1] - y3;
    xl[2] = xyz[3 * size_t(c.y)] - x3; xl[3] = xyz[3 * size_t(c.y) + 1] - y3;
    xl[4] = xyz[3 * size_t(c.w)] - x3; xl[5] = xyz[3 * size_t(c.w) + 1] - y3;
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
    if (ip == 0) {
      const double4 prm = *reinterpret_cast<const double4*>(params + 4 * size_t(e));
      const fem::PlateMat m = fem::plate_material(prm.x, prm.y, prm.z, prm.w);
      fem::PlateShear sv;
      fem::plate_shear_vectors(xl, sv);
      double* ed = edat + le * PF_EDAT;
      ed[0] = m.cm00; ed[1] = m.cm01; ed[2] = m.cm22; ed[3] = m.cb00; ed[4] = m.cb01; ed[5] = m.cb22; ed[6] = m.cs;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) { ed[8 + 3 * a + i] = sv.v[a][i]; ed[20 + 3 * a + i] = sv.u[a][i]; }
    }
  }
  if (t == 0) {
    uint32_t off = 0;
    for (uint32_t o = 0; o < nown; ++o) {
      s_ooff[o] = off;
      off += 13 * ((s_info[o] >> 8) & 255u) + 1;
    }
    s_ooff[nown] = off;
  }
  __syncthreads();

  // ---- step 2: owned rows folded over the four points
  double row[PF_SLOT - 2];
  if (t < int(nrows)) {
    const uint32_t ref = rows[r0 + t];
    const int re = int(ref & 255u), p = int(ref >> 8);
    double acc[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) acc[k] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double2* gi = geo2 + (re * 4 + q) * (PF_GEO / 2);
      double gv[18];
#pragma unroll
      for (int h = 0; h < 9; ++h) { const double2 v = gi[h]; gv[2 * h] = v.x; gv[2 * h + 1] = v.y; }
      fem::PlatePoint g;
      g.w = gv[0];
#pragma unroll
      for (int a = 0; a < 4; ++a) { g.dx[a] = gv[1 + a]; g.dy[a] = gv[5 + a]; g.al[a] = gv[9 + a]; g.be[a] = gv[13 + a]; }
      const double wdx = g.w * fem::sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p);
      const double wdy = g.w * fem::sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
      const double wal = g.w * fem::sel4(g.al[0], g.al[1], g.al[2], g.al[3], p);
      const double wbe = g.w * fem::sel4(g.be[0], g.be[1], g.be[2], g.be[3], p);
      fem::plate_row_accumulate(g, wdx, wdy, wal, wbe, acc);
    }
    const double* ed = edat + re * PF_EDAT;
    fem::PlateMat m;
    m.cm00 = ed[0]; m.cm01 = ed[1]; m.cm22 = ed[2]; m.cb00 = ed[3]; m.cb01 = ed[4]; m.cb22 = ed[5]; m.cs = ed[6];
    fem::PlateShear sv;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) { sv.v[a][i] = ed[8 + 3 * a + i]; sv.u[a][i] = ed[20 + 3 * a + i]; }
    double vp[3], up[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { vp[i] = m.cs * ed[8 + 3 * p + i]; up[i] = m.cs * ed[20 + 3 * p + i]; }
    fem::plate_row_finish<PF_BS>(acc, m, vp, up, sv, row);
    stage_p[t] = p;
  }
  __syncthreads();

  // ---- step 3: stage the rows (over the step-1 data)
  if (t < int(nrows)) {
#pragma unroll
    for (int k = 0; k < (PF_SLOT - 2) / 2; ++k) lds2[t * (PF_SLOT / 2) + k] = make_double2(row[2 * k], row[2 * k + 1]);
  }
  __syncthreads();

  // ---- step 4: one lane per owned node-pair block: contributors summed in
  //      element-insertion order (registers), then placed in CSR order (LDS)
  constexpr int MAXB = (PF_NODES * 16 + PF_THREADS - 1) / PF_THREADS;
  double bsum[MAXB][13];
  int bpos[MAXB];
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    const uint32_t bi = uint32_t(t + it * PF_THREADS);
    bpos[it] = -1;
    if (bi >= nblk) continue;
    const uint32_t d = blk[b0 + bi];
    const int n = int((d >> 23) & 63u);
    const uint32_t info = s_info[n];
    const int deg = int(info & 255u), sb = int(info >> 24);
    const bool self = (d >> 29) & 1u;
    const int cnt = self ? deg : int(d & 7u);
    for (int m = 0; m < cnt; ++m) {
      int r, q;
      if (self) { r = m; q = stage_p[sb + m]; }
      else { r = int((d >> (3 + 5 * m)) & 7u); q = int((d >> (6 + 5 * m)) & 3u); }
      const double2* src = lds2 + ((sb + r) * PF_SLOT + PF_BS * q) / 2;
      double v[14];
#pragma unroll
      for (int h = 0; h < 7; ++h) { const double2 w = src[h]; v[2 * h] = w.x; v[2 * h + 1] = w.y; }
#pragma unroll
      for (int j = 0; j < 13; ++j) bsum[it][j] = (m == 0) ? v[j] : bsum[it][j] + v[j];
    }
    bpos[it] = int(bi);
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXB; ++it) {
    if (bpos[it] < 0) continue;
    const uint32_t bi = uint32_t(bpos[it]);
    const uint32_t d = blk[b0 + bi];
    const int n = int((d >> 23) & 63u);
    const uint32_t info = s_info[n];
    const int deg = int(info & 255u), nb = int((info >> 8) & 255u);
    const int k = int(bi - s_boff[n]);
    double* o = lds + s_ooff[n];
    o[2 * k] = bsum[it][0];
    o[2 * k + 1] = bsum[it][1];
    o[2 * nb + 2 * k] = bsum[it][2];
    o[2 * nb + 2 * k + 1] = bsum[it][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      o[4 * nb + 3 * k + j] = bsum[it][4 + j];
      o[7 * nb + 3 * k + j] = bsum[it][7 + j];
      o[10 * nb + 3 * k + j] = bsum[it][10 + j];
    }
    if ((d >> 29) & 1u) {
      // drilling row: KROT6 of every element, in insertion order
      double dr = fem::KROT6;
      for (int r = 1; r < deg; ++r) dr += fem::KROT6;
      o[13 * nb] = dr;
    }
  }
  __syncthreads();

  // ---- step 5: coalesced write-out of the owned CSR segments
  const int lane = t & 63, wave = t >> 6;
  for (int o = wave; o < int(nown); o += PF_THREADS / 64) {
    const int len = int(s_ooff[o + 1] - s_ooff[o]);
    const double* src = lds + s_ooff[o];
    double* dst = val + s_base[o];
    for (int i = lane; i < len; i += 64) dst[i] = src[i];
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
  if (!(e = dupload(d->pf_elem_ptr, pm.patch_elem_ptr.data(), pm.patch_elem_ptr.size())).empty()) return e;
  if (!(e = dupload(d->pf_elems, pm.elems.data(), pm.elems.size())).empty()) return e;
  if (!(e = dupload(d->pf_row_ptr, pm.patch_row_ptr.data(), pm.patch_row_ptr.size())).empty()) return e;
  return dupload(d->pf_rows, pm.rows.data(), pm.rows.size());
}

std::string assemble_plate_fused(Device* d, double* val, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipEventRecord(d->ev0, s));
  if (d->n_patches > 0) {
    plate_fused_kernel<<<d->n_patches, PF_THREADS, 0, s>>>(d->xyz, d->conn[PLATE], d->params[PLATE], d->pf_node_ptr,
                                                         d->pf_node, d->pf_info, d->pf_blk_ptr, d->pf_blk,
                                                         d->pf_elem_ptr, d->pf_elems, d->pf_row_ptr, d->pf_rows,
                                                         d->row_ptr, val);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(d->ev1, s));
  d->timed = true;
  return "";
}

}  // namespace femdev
