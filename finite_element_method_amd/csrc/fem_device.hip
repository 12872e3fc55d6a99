// fem_device.hip -- gfx950 device runtime of the assembly path.
//
// HBM layout (all FP64 values, 32-bit indices):
//   xyz      N x 3 node coordinates (AoS: one 24-byte gather per node)
//   conn[t]  E_t x nodes_per node indices, params[t] E_t x params_per
//   R[t]     E_t x 9 rotation frames computed by the setup kernels
//            (Truss/Beam/Plate::create of the crate)
//   row_ptr  6N + 1 int64, col_idx nnz int32, val nnz double  (CSR)
//
// Value paths (both deterministic, no atomics; every CSR slot is written once):
//   generic      : items (element, node pair) write their structural entries
//                  into a contribution buffer; slots then sum their
//                  contributions in element-insertion order (a sorted,
//                  segmented reduction through a precomputed slot map).
//   plate_fused  : one workgroup per patch of owned nodes; element kinematics
//                  are staged in LDS and each owned CSR row block is summed in
//                  registers and written once (see plate_fused_kernel).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

using fem::NODE_DOF;

Device* create_device(std::string& err) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    err = "no HIP device available: the assembly path runs on the GPU only";
    return nullptr;
  }
  auto* d = new Device();
  hipGetDevice(&d->dev);
  if (hipEventCreate(&d->ev0) != hipSuccess || hipEventCreate(&d->ev1) != hipSuccess ||
      hipEventCreate(&d->ev2) != hipSuccess || hipEventCreate(&d->ev3) != hipSuccess ||
      hipEventCreate(&d->ev4) != hipSuccess || hipEventCreate(&d->ev5) != hipSuccess) {
    err = "hipEventCreate failed";
    delete d;
    return nullptr;
  }
  return d;
}

void destroy_device(Device* d) {
  if (!d) return;
  auto f = [](auto*& p) { if (p) { hipFree(p); p = nullptr; } };
  f(d->xyz);
  for (int t = 0; t < 3; ++t) { f(d->conn[t]); f(d->params[t]); f(d->R[t]); f(d->rzero[t]); f(d->eclass[t]); f(d->ebase[t]); }
  f(d->derived); f(d->truss_dir); f(d->row_ptr); f(d->col_idx); f(d->val); f(d->class_mask); f(d->class_off); f(d->contrib);
  f(d->slot_ptr); f(d->slot_src); f(d->pf_hdr); f(d->pf_conn); f(d->pf_eid); f(d->pf_rows); f(d->pf_oinfo);
  f(d->pf_oboff); f(d->pf_oooff); f(d->pf_blk); f(d->pf_obase);
  for (int t = 0; t < 3; ++t) f(d->adj_bid[t]);
  for (int t = 0; t < 2; ++t) f(d->bo_bid[t]);
  f(d->bo_bmask); f(d->bo_boff); f(d->bo_inc_ptr); f(d->bo_inc); f(d->bo_ranges);
  f(d->bo_slot_inc); f(d->bo_slot_ptr); f(d->bo_slot);
  f(d->adj_bmask); f(d->adj_boff); f(d->adj_inc_ptr); f(d->adj_xpos); f(d->adj_xs); f(d->adj_kpack); f(d->adj_eorder);
  f(d->sep_cons); f(d->sep_pack); f(d->sep_cls); f(d->sep_tmp); f(d->sep_disp); f(d->sep_force); f(d->sep_ra); f(d->sep_ub);
  f(d->sep_v); f(d->sep_flag); f(d->sep_pos); f(d->sep_key); f(d->sep_cnt); f(d->sep_off); f(d->sep_aa); f(d->sep_bb);
  f(d->sep_rp); f(d->sep_sum); f(d->sep_ci); f(d->sep_dpos);
  f(d->pb_hdr); f(d->pb_conn); f(d->pb_eid); f(d->pb_blk);
  f(d->pl_nodes); f(d->pl_lanes); f(d->pl_lane_set); f(d->pl_geo);
  for (int k = 0; k < 6; ++k) f(d->sol_vec[k]);
  f(d->sol_rr); f(d->sol_rb); f(d->sol_part); f(d->sol_s); f(d->sol_diag); f(d->sol_inv); f(d->sol_flag);
  f(d->sol_bid); f(d->sol_starts); f(d->sol_tmp); f(d->u_glob);
  if (d->sol_host) { (void)hipHostFree(d->sol_host); d->sol_host = nullptr; }
  for (hipEvent_t ev : {d->ev0, d->ev1, d->ev2, d->ev3, d->ev4, d->ev5})
    if (ev) hipEventDestroy(ev);
  delete d;
}

std::string upload_nodes(Device* d, const double* xyz, size_t n) {
  d->n_nodes = n;
  return dupload(d->xyz, xyz, 3 * n);
}

std::string upload_elements(Device* d, int type, size_t n, const uint32_t* conn_idx, const double* params) {
  d->n_elem[type] = n;
  auto e = dupload(d->conn[type], conn_idx, n * NODES_PER[type]);
  if (!e.empty()) return e;
  return dupload(d->params[type], params, n * PARAMS_PER[type]);
}

// ------------------------------------------------------------------ setup kernels
__device__ inline uint16_t zero_bits(const double R[9]) {
  uint16_t b = 0;
  for (int k = 0; k < 9; ++k) b |= uint16_t(R[k] != 0.0) << k;
  return b;
}

// Truss::create (truss.rs:230-255): direction frame
__global__ void truss_setup_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ xyz, size_t n,
                                   double abs_tol, double* __restrict__ Rout, double* __restrict__ dir,
                                   uint16_t* __restrict__ rz) {
  size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint32_t a = conn[2 * e], b = conn[2 * e + 1];
  double d[3] = {xyz[3 * b] - xyz[3 * a], xyz[3 * b + 1] - xyz[3 * a + 1], xyz[3 * b + 2] - xyz[3 * a + 2]};
  double R[9];
  bool ok = fem::truss_rotation(d, abs_tol, R);
  if (!ok) for (int k = 0; k < 9; ++k) R[k] = 0.0;
  for (int k = 0; k < 9; ++k) Rout[9 * e + k] = R[k];
  for (int k = 0; k < 3; ++k) dir[3 * e + k] = R[k];  // compact copy of the row the assembly reads
  rz[e] = ok ? zero_bits(R) : uint16_t(0x8000);
}

// Beam::create (beam.rs:687-745): principal moments and the rolled frame
__global__ void beam_setup_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ xyz,
                                  const double* __restrict__ params, size_t n, double rel_tol, double abs_tol,
                                  double* __restrict__ Rout, double* __restrict__ derived,
                                  uint16_t* __restrict__ rz) {
  size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* p = params + 11 * e;
  uint32_t a = conn[2 * e], b = conn[2 * e + 1];
  double d[3] = {xyz[3 * b] - xyz[3 * a], xyz[3 * b + 1] - xyz[3 * a + 1], xyz[3 * b + 2] - xyz[3 * a + 2]};
  double i11p, i22p, angle;
  fem::beam_principal(p[3], p[4], p[5], rel_tol, i11p, i22p, angle);
  double v1[3] = {p[8], p[9], p[10]};
  double R[9];
  bool ok = fem::beam_rotation(d, v1, angle, abs_tol, R);
  if (!ok) for (int k = 0; k < 9; ++k) R[k] = 0.0;
  for (int k = 0; k < 9; ++k) Rout[9 * e + k] = R[k];
  derived[2 * e] = i11p;
  derived[2 * e + 1] = i22p;
  rz[e] = ok ? zero_bits(R) : uint16_t(0x8000);
}

// Plate::create (plate.rs:1050-1065): element frame from nodes 2, 3, 4
__global__ void plate_setup_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ xyz, size_t n,
                                   double abs_tol, double* __restrict__ Rout, uint16_t* __restrict__ rz) {
  size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t* c = conn + 4 * e;
  double p2[3] = {xyz[3 * c[1]], xyz[3 * c[1] + 1], xyz[3 * c[1] + 2]};
  double p3[3] = {xyz[3 * c[2]], xyz[3 * c[2] + 1], xyz[3 * c[2] + 2]};
  double p4[3] = {xyz[3 * c[3]], xyz[3 * c[3] + 1], xyz[3 * c[3] + 2]};
  double R[9];
  bool ok = fem::quad_rotation(p2, p3, p4, abs_tol, R);
  if (!ok) for (int k = 0; k < 9; ++k) R[k] = 0.0;
  for (int k = 0; k < 9; ++k) Rout[9 * e + k] = R[k];
  bool ident = R[0] == 1.0 && R[4] == 1.0 && R[8] == 1.0 && R[1] == 0.0 && R[2] == 0.0 && R[3] == 0.0 &&
               R[5] == 0.0 && R[6] == 0.0 && R[7] == 0.0;
  rz[e] = ok ? uint16_t(zero_bits(R) | (ident ? 0x4000 : 0)) : uint16_t(0x8000);
}

std::string run_setup(Device* d, int type, size_t n, double rel_tol, double abs_tol, std::vector<uint16_t>& rzero) {
  d->rel_tol = rel_tol;
  d->abs_tol = abs_tol;
  rzero.assign(n, 0);
  if (n == 0) return "";
  auto e = dalloc(d->R[type], 9 * n);
  if (!e.empty()) return e;
  e = dalloc(d->rzero[type], n);
  if (!e.empty()) return e;
  unsigned blocks = unsigned((n + 255) / 256);
  if (type == TRUSS) {
    e = dalloc(d->truss_dir, 3 * n);
    if (!e.empty()) return e;
    truss_setup_kernel<<<blocks, 256>>>(d->conn[0], d->xyz, n, abs_tol, d->R[0], d->truss_dir, d->rzero[0]);
  } else if (type == BEAM) {
    e = dalloc(d->derived, 2 * n);
    if (!e.empty()) return e;
    beam_setup_kernel<<<blocks, 256>>>(d->conn[1], d->xyz, d->params[1], n, rel_tol, abs_tol, d->R[1], d->derived,
                                       d->rzero[1]);
  } else {
    plate_setup_kernel<<<blocks, 256>>>(d->conn[2], d->xyz, n, abs_tol, d->R[2], d->rzero[2]);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(rzero.data(), d->rzero[type], n * sizeof(uint16_t), hipMemcpyDeviceToHost));
  return "";
}

std::string replace_params(Device* d, int type, const double* src, double* host_copy, void* stream) {
  const size_t n = d->n_elem[type] * PARAMS_PER[type];
  if (!n) return "";
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(d->params[type], src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (type == PLATE) d->pl_geo_dirty = true;
  else d->bo_slot_dirty = true;
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipMemcpy(host_copy, d->params[type], n * sizeof(double), hipMemcpyDeviceToHost));
  return "";
}

std::string replace_nodes(Device* d, const double* src, double* host_copy, void* stream) {
  const size_t n = 3 * d->n_nodes;
  if (!n) return "";
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(d->xyz, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  d->pl_geo_dirty = true;
  d->bo_slot_dirty = true;
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipMemcpy(host_copy, d->xyz, n * sizeof(double), hipMemcpyDeviceToHost));
  return "";
}

std::string get_rotation(Device* d, int type, size_t pos, double out[9]) {
  if (!d->R[type] || pos >= d->n_elem[type]) return "rotation frames are not computed";
  HIP_TRY(hipMemcpy(out, d->R[type] + 9 * pos, 9 * sizeof(double), hipMemcpyDeviceToHost));
  return "";
}

std::string upload_pattern(Device* d, int64_t n_rows, const int64_t* rp, int64_t nnz, const int32_t* ci) {
  d->n_rows = n_rows;
  d->nnz = nnz;
  d->sep_dpos_valid = false;
  d->sep_done = false;
  auto e = dupload(d->row_ptr, rp, size_t(n_rows + 1));
  if (!e.empty()) return e;
  e = dupload(d->col_idx, ci, size_t(nnz));
  if (!e.empty()) return e;
  return dalloc(d->val, size_t(nnz));
}

// ------------------------------------------------------------------ element blocks
// Global-frame 6x6 block (p, q) of a plate's K_e = T^T k T
// (plate.rs:513-724 local k, methods_for_plate_data_handle.rs:104-112 rotation).
__device__ inline void plate_block(const double* __restrict__ xyz, const uint32_t* c, const double* prm,
                                   const double R[9], int p, int q, double kb[36]) {
  double X[4][3];
  for (int a = 0; a < 4; ++a)
    for (int k = 0; k < 3; ++k) X[a][k] = xyz[3 * c[a] + k];
  double xl[6];
  const int others[3] = {0, 1, 3};
  for (int m = 0; m < 3; ++m) {
    double dd[3] = {X[others[m]][0] - X[2][0], X[others[m]][1] - X[2][1], X[others[m]][2] - X[2][2]};
    xl[2 * m] = (R[0] * dd[0] + R[1] * dd[1]) + R[2] * dd[2];
    xl[2 * m + 1] = (R[3] * dd[0] + R[4] * dd[1]) + R[5] * dd[2];
  }
  fem::PlateMat mat = fem::plate_material(prm[0], prm[1], prm[2], prm[3]);
  for (int i = 0; i < 36; ++i) kb[i] = 0.0;
  const double g = fem::plate_gauss();
  const double rs[4][2] = {{g, g}, {-g, g}, {-g, -g}, {g, -g}};
  for (int ip = 0; ip < 4; ++ip) {
    fem::PlateIP gi;
    fem::plate_ip(xl, rs[ip][0], rs[ip][1], gi);
    fem::plate_block_add_ip(gi, mat, p, q, kb);
  }
  if (p == q) kb[35] += fem::KROT6;
  fem::rotate_block(kb, R);
}

// Global-frame 6x6 block (p, q) of a beam's K_e (beam.rs:495-605, one point r = 0, alpha = 2)
__device__ inline void beam_block(const double* __restrict__ xyz, const uint32_t* c, const double* prm,
                                  const double* der, const double R[9], int p, int q, double kb[36]) {
  double d[3] = {xyz[3 * c[1]] - xyz[3 * c[0]], xyz[3 * c[1] + 1] - xyz[3 * c[0] + 1],
                 xyz[3 * c[1] + 2] - xyz[3 * c[0] + 2]};
  double L = fem::norm3(d[0], d[1], d[2]);
  double J = fem::bar_jacobian(L);
  double inv_j = 1.0 / J, det_j = J, alpha = 2.0;
  const double E = prm[0], nu = prm[1], A = prm[2], it = prm[6], k = prm[7];
  double G = E / (2.0 * (1.0 + nu));
  double scale[6] = {A * E * det_j * alpha,  G * A * k * det_j * alpha, G * A * k * det_j * alpha,
                     G * it * det_j * alpha, E * der[1] * det_j * alpha, E * der[0] * det_j * alpha};
  double bp = (p == 0 ? 0.0 - 0.5 : 0.0 + 0.5) * inv_j, bq = (q == 0 ? 0.0 - 0.5 : 0.0 + 0.5) * inv_j;
  // rank-one terms b_t b_t^T restricted to nodes (p, q); dof order u v w tu tv tw
  double vp[6][6] = {{bp, 0, 0, 0, 0, 0},     {0, bp, 0, 0, 0, -0.5}, {0, 0, bp, 0, -0.5, 0},
                     {0, 0, 0, bp, 0, 0},     {0, 0, 0, 0, bp, 0},    {0, 0, 0, 0, 0, bp}};
  double vq[6][6] = {{bq, 0, 0, 0, 0, 0},     {0, bq, 0, 0, 0, -0.5}, {0, 0, bq, 0, -0.5, 0},
                     {0, 0, 0, bq, 0, 0},     {0, 0, 0, 0, bq, 0},    {0, 0, 0, 0, 0, bq}};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double acc = 0.0;
      for (int t = 0; t < 6; ++t) acc += (vp[t][i] * vq[t][j]) * scale[t];
      kb[i * 6 + j] = acc;
    }
  fem::rotate_block(kb, R);
}

// Global-frame translational 3x3 block (p, q) of a truss (truss.rs:132-189), in a 6x6 layout
__device__ inline void truss_block(const double* __restrict__ xyz, const uint32_t* c, const double* prm,
                                   const double R[9], int p, int q, double kb[36]) {
  double d[3] = {xyz[3 * c[1]] - xyz[3 * c[0]], xyz[3 * c[1] + 1] - xyz[3 * c[0] + 1],
                 xyz[3 * c[1] + 2] - xyz[3 * c[0] + 2]};
  double L = fem::norm3(d[0], d[1], d[2]);
  double J = fem::bar_jacobian(L);
  double inv_j = 1.0 / J;
  const double E = prm[0], A = prm[1], A2 = prm[2];
  double r = 0.0;
  double area = (A2 == A2) ? (A2 - A) / 2.0 * r + A - (A2 - A) / 2.0 * -1.0 : A;
  double scale = area * E * J * 2.0;
  double bp = (p == 0 ? -0.5 : 0.5) * inv_j, bq = (q == 0 ? -0.5 : 0.5) * inv_j;
  double k00 = (bp * bq) * scale;
  for (int i = 0; i < 36; ++i) kb[i] = 0.0;
  // R^T diag(k00, 0, 0) R = k00 r0 r0^T
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) kb[i * 6 + j] = (R[i] * k00) * R[j];
}

// ------------------------------------------------------------------ generic path
template <int TYPE>
__global__ __launch_bounds__(512) void generic_items_kernel(const double* __restrict__ xyz, const uint32_t* __restrict__ conn,
                                     const double* __restrict__ params, const double* __restrict__ Rall,
                                     const double* __restrict__ derived, const uint16_t* __restrict__ eclass,
                                     const int64_t* __restrict__ ebase, const uint64_t* __restrict__ cmask,
                                     const uint32_t* __restrict__ coff, size_t n_items,
                                     double* __restrict__ contrib) {
  constexpr int NN = TYPE == PLATE ? 4 : 2;
  constexpr int NP = TYPE == PLATE ? 4 : (TYPE == BEAM ? 11 : 3);
  size_t it = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (it >= n_items) return;
  size_t e = it / (NN * NN);
  int pq = int(it % (NN * NN));
  int p = pq / NN, q = pq % NN;
  const uint32_t* c = conn + NN * e;
  double R[9];
  for (int k = 0; k < 9; ++k) R[k] = Rall[9 * e + k];
  double kb[36];
  if (TYPE == PLATE) plate_block(xyz, c, params + NP * e, R, p, q, kb);
  else if (TYPE == BEAM) beam_block(xyz, c, params + NP * e, derived + 2 * e, R, p, q, kb);
  else truss_block(xyz, c, params + NP * e, R, p, q, kb);
  unsigned cls = eclass[e];
  uint64_t mask = cmask[cls * 16 + pq];
  double* out = contrib + ebase[e] + coff[cls * 16 + pq];
  int k = 0;
  for (int b = 0; b < 36; ++b)
    if ((mask >> b) & 1ull) out[k++] = kb[b];
}

// val[v] = sum of its contributions in element-insertion order
__global__ void slot_gather_kernel(const int64_t* __restrict__ slot_ptr, const uint32_t* __restrict__ slot_src,
                                   const double* __restrict__ contrib, int64_t nnz, double* __restrict__ val) {
  int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nnz) return;
  int64_t b = slot_ptr[v], e = slot_ptr[v + 1];
  double acc = contrib[slot_src[b]];
  for (int64_t k = b + 1; k < e; ++k) acc += contrib[slot_src[k]];
  val[v] = acc;
}

std::string upload_generic(Device* d, const GenericMaps& g) {
  auto e = dupload(d->class_mask, g.class_mask.data(), g.class_mask.size());
  if (!e.empty()) return e;
  e = dupload(d->class_off, g.class_off.data(), g.class_off.size());
  if (!e.empty()) return e;
  for (int t = 0; t < 3; ++t) {
    e = dupload(d->eclass[t], g.elem_class[t].data(), g.elem_class[t].size());
    if (!e.empty()) return e;
    e = dupload(d->ebase[t], g.elem_base[t].data(), g.elem_base[t].size());
    if (!e.empty()) return e;
  }
  d->n_contrib = g.n_contrib;
  e = dalloc(d->contrib, size_t(g.n_contrib));
  if (!e.empty()) return e;
  e = dupload(d->slot_ptr, g.slot_ptr.data(), g.slot_ptr.size());
  if (!e.empty()) return e;
  return dupload(d->slot_src, g.slot_src.data(), g.slot_src.size());
}

void set_generic_block(Device* d, int threads) { d->gen_block = threads; }

std::string assemble_generic(Device* d, double* val, void* stream, bool timed) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned bs = unsigned(d->gen_block);
  if (timed) HIP_TRY(hipEventRecord(d->ev0, s));
  for (int t = 0; t < 3; ++t) {
    size_t n = d->n_elem[t];
    if (!n) continue;
    size_t items = n * (t == PLATE ? 16 : 4);
    unsigned blocks = unsigned((items + bs - 1) / bs);
    if (t == PLATE)
      generic_items_kernel<PLATE><<<blocks, bs, 0, s>>>(d->xyz, d->conn[t], d->params[t], d->R[t], d->derived,
                                                         d->eclass[t], d->ebase[t], d->class_mask, d->class_off,
                                                         items, d->contrib);
    else if (t == BEAM)
      generic_items_kernel<BEAM><<<blocks, bs, 0, s>>>(d->xyz, d->conn[t], d->params[t], d->R[t], d->derived,
                                                        d->eclass[t], d->ebase[t], d->class_mask, d->class_off,
                                                        items, d->contrib);
    else
      generic_items_kernel<TRUSS><<<blocks, bs, 0, s>>>(d->xyz, d->conn[t], d->params[t], d->R[t], d->derived,
                                                         d->eclass[t], d->ebase[t], d->class_mask, d->class_off,
                                                         items, d->contrib);
    HIP_TRY(hipGetLastError());
  }
  if (d->nnz > 0) {
    unsigned blocks = unsigned((d->nnz + bs - 1) / bs);
    slot_gather_kernel<<<blocks, bs, 0, s>>>(d->slot_ptr, d->slot_src, d->contrib, d->nnz, val);
    HIP_TRY(hipGetLastError());
  }
  if (timed) {
    HIP_TRY(hipEventRecord(d->ev1, s));
    d->timed = true;
  }
  return "";
}

double last_ms(Device* d) {
  if (!d->timed) return -1.0;
  if (hipEventSynchronize(d->ev1) != hipSuccess) return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, d->ev0, d->ev1) != hipSuccess) return -1.0;
  return double(ms);
}

double* values(Device* d) { return d->val; }
const int64_t* row_ptr(Device* d) { return d->row_ptr; }
const int32_t* col_idx(Device* d) { return d->col_idx; }

std::string copy_csr_to_host(Device* d, int64_t n_rows, int64_t nnz, int64_t* rp, int32_t* ci, double* v) {
  HIP_TRY(hipDeviceSynchronize());
  if (rp && n_rows >= 0) HIP_TRY(hipMemcpy(rp, d->row_ptr, size_t(n_rows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (ci && nnz) HIP_TRY(hipMemcpy(ci, d->col_idx, size_t(nnz) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (v && nnz) HIP_TRY(hipMemcpy(v, d->val, size_t(nnz) * sizeof(double), hipMemcpyDeviceToHost));
  return "";
}

}  // namespace femdev
