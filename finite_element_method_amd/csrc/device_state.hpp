// device_state.hpp -- device-side state of one model, shared by the .hip files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace femdev {

struct Device {
  hipStream_t stream = nullptr;
  int dev = 0;
  double* xyz = nullptr;
  size_t n_nodes = 0;
  uint32_t* conn[3] = {nullptr, nullptr, nullptr};
  double* params[3] = {nullptr, nullptr, nullptr};
  double* R[3] = {nullptr, nullptr, nullptr};
  double* truss_dir = nullptr;  // trusses: R's first row, 3 doubles each (the only row the stiffness reads)
  double* derived = nullptr;  // beams: i11p, i22p
  uint16_t* rzero[3] = {nullptr, nullptr, nullptr};
  size_t n_elem[3] = {0, 0, 0};
  // CSR
  int64_t* row_ptr = nullptr;
  int32_t* col_idx = nullptr;
  double* val = nullptr;
  int64_t n_rows = 0, nnz = 0;
  // generic maps
  uint64_t* class_mask = nullptr;
  uint32_t* class_off = nullptr;
  uint16_t* eclass[3] = {nullptr, nullptr, nullptr};
  int64_t* ebase[3] = {nullptr, nullptr, nullptr};
  double* contrib = nullptr;
  int64_t n_contrib = 0;
  int64_t* slot_ptr = nullptr;
  uint32_t* slot_src = nullptr;
  // plate fused maps
  int n_patches = 0, pf_grid = 0, gen_block = 256;
  uint32_t *pf_hdr = nullptr, *pf_conn = nullptr, *pf_eid = nullptr, *pf_rows = nullptr, *pf_oinfo = nullptr,
           *pf_oboff = nullptr, *pf_oooff = nullptr, *pf_blk = nullptr;
  int64_t* pf_obase = nullptr;
  // plate block-lane maps
  int pb_patches = 0, pb_grid = 0;
  uint32_t *pb_hdr = nullptr, *pb_conn = nullptr, *pb_eid = nullptr, *pb_blk = nullptr;
  // plate balanced-lane maps (the element lists share pb_conn / pb_eid)
  int pl_waves = 0;
  uint32_t *pl_nodes = nullptr, *pl_lanes = nullptr, *pl_lane_set = nullptr;
  double* pl_geo = nullptr;   // per wave element slot: local coordinates and parameters
  bool pl_geo_dirty = false;  // node coordinates or plate parameters changed since pl_geo was filled
  // every plate has the same E, nu and k (bitwise): the slots hold t only and
  // the three values go to the kernel as arguments
  bool pl_uniform = false;
  double pl_mat[3] = {0.0, 0.0, 0.0};
  // bar owner maps
  uint32_t* bo_bid[2] = {nullptr, nullptr};
  uint64_t* bo_bmask = nullptr;
  uint16_t* bo_boff = nullptr;
  uint32_t *bo_inc_ptr = nullptr, *bo_inc = nullptr, *bo_ranges = nullptr;
  size_t bo_n_ranges = 0, bo_n_truss_ranges = 0;  // bar-kernel ranges, then truss-kernel ranges
  uint32_t *bo_slot_inc = nullptr, *bo_slot_ptr = nullptr;
  double* bo_slot = nullptr;   // staged entry constants of the bar-kernel ranges
  size_t bo_n_slots = 0;
  bool bo_slot_dirty = false;  // coordinates or bar parameters changed since bo_slot was filled
  // adjoint maps
  uint32_t* adj_bid[3] = {nullptr, nullptr, nullptr};
  uint64_t* adj_bmask = nullptr;
  uint16_t* adj_boff = nullptr;
  uint32_t *adj_inc_ptr = nullptr, *adj_xpos = nullptr, *adj_kpack = nullptr, *adj_eorder = nullptr;
  double* adj_xs = nullptr;
  int64_t adj_xoff[3] = {0, 0, 0};
  int64_t adj_nxs = 0;
  bool adj_flat = false;
  double rel_tol = 0.0, abs_tol = 0.0;
  // boundary conditions per global dof and the sparse separation; the four
  // matrices (K_aa, K_ab, K_ba, K_bb) share sep_rp / sep_ci / sep_v
  uint8_t *sep_cons = nullptr, *sep_cls = nullptr, *sep_tmp = nullptr;
  double *sep_disp = nullptr, *sep_force = nullptr, *sep_ra = nullptr, *sep_ub = nullptr, *sep_v = nullptr;
  uint64_t *sep_flag = nullptr, *sep_pos = nullptr;
  int64_t *sep_key = nullptr, *sep_cnt = nullptr, *sep_off = nullptr, *sep_aa = nullptr, *sep_bb = nullptr;
  int64_t *sep_rp = nullptr, *sep_sum = nullptr;
  int32_t* sep_ci = nullptr;
  int64_t* sep_dpos = nullptr;  // diagonal entry of each row (pattern-only)
  uint32_t* sep_pack = nullptr;  // per dof: class << 30 | its index among the a (or b) dofs
  bool sep_dpos_valid = false;
  size_t sep_cap[20] = {};
  int64_t sep_bc_rows = -1;
  int64_t sep_rows[4] = {0, 0, 0, 0}, sep_nnz[4] = {0, 0, 0, 0}, sep_base[4] = {0, 0, 0, 0};
  bool sep_done = false;
  double sep_ms = 0.0;
  // sparse global analysis (solve.hip): b, x = u_a, r, z, p, Ap
  double* sol_vec[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double *sol_rr = nullptr, *sol_rb = nullptr, *sol_part = nullptr, *sol_s = nullptr, *sol_host = nullptr;
  double *sol_diag = nullptr, *sol_inv = nullptr;
  uint32_t *sol_flag = nullptr, *sol_bid = nullptr;
  int32_t* sol_starts = nullptr;
  uint8_t* sol_tmp = nullptr;
  size_t sol_cap[16] = {};
  int64_t sol_na = 0, sol_nb = 0;
  bool sol_done = false;
  double sol_ms = 0.0;
  double* u_glob = nullptr;  // global displacements of the last solve (6 per node)
  size_t ucap = 0;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr, ev5 = nullptr;
  bool timed = false, timed_bwd = false;
};

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return std::string(#expr) + ": " + hipGetErrorString(_e);  \
  } while (0)

template <class T>
inline std::string dalloc(T*& p, size_t n) {
  if (p) { (void)hipFree(p); p = nullptr; }
  if (n == 0) return "";
  HIP_TRY(hipMalloc(&p, n * sizeof(T)));
  return "";
}
template <class T>
inline std::string dupload(T*& p, const T* src, size_t n) {
  auto e = dalloc(p, n);
  if (!e.empty()) return e;
  if (n) HIP_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return "";
}

}  // namespace femdev
