// fem_device.hpp -- interface between the host model (fem_host.cpp) and the
// device runtime (fem_device.hip).  Plain pointers and sizes only.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace femdev {

struct Device;  // opaque: device buffers, maps and streams of one model

Device* create_device(std::string& err);
void destroy_device(Device* d);

// element type ids (match fem_element_type)
constexpr int TRUSS = 0, BEAM = 1, PLATE = 2;
constexpr int NODES_PER[3] = {2, 2, 4};
constexpr int PARAMS_PER[3] = {3, 11, 4};

// Uploads node coordinates (n x 3) and one element type's connectivity (node
// indices, n x nodes_per) and parameters (n x params_per).
std::string upload_nodes(Device* d, const double* xyz, size_t n);
std::string upload_elements(Device* d, int type, size_t n, const uint32_t* conn_idx, const double* params);

// Runs the per-element setup on the device (rotation frames; beam principal
// moments) and returns the zero pattern of each element's 3x3 rotation (bit k
// set when R[k] != 0) in `rzero` (n entries).  A set bit 15 flags an element
// whose frame could not be formed (zero-length vector).
std::string run_setup(Device* d, int type, size_t n, double rel_tol, double abs_tol, std::vector<uint16_t>& rzero);

// Rotation matrix of element `pos` (insertion position within its type).
std::string get_rotation(Device* d, int type, size_t pos, double out[9]);

// CSR pattern upload (row_ptr: n_rows + 1, col_idx: nnz).
std::string upload_pattern(Device* d, int64_t n_rows, const int64_t* row_ptr, int64_t nnz, const int32_t* col_idx);

// Generic value path: element items write their structural entries into a
// contribution buffer, slots gather them in element-insertion order.
struct GenericMaps {
  // per element class: per pair mask (bit 6*i+j of the node-pair block) and the
  // offset of the pair's first entry within the element record
  std::vector<uint64_t> class_mask;  // n_class * 16
  std::vector<uint32_t> class_off;   // n_class * 16
  // per type: element class id and record base in the contribution buffer
  std::vector<uint16_t> elem_class[3];
  std::vector<int64_t> elem_base[3];
  int64_t n_contrib = 0;
  std::vector<int64_t> slot_ptr;   // nnz + 1
  std::vector<uint32_t> slot_src;  // n_contrib
};
std::string upload_generic(Device* d, const GenericMaps& g);

// Plate fused path (flat plates whose frame is exactly the identity): one
// workgroup per patch of owned nodes; see plate_fused.hip.  Limits per patch:
constexpr int PF_MAX_ELEMS = 64;   // elements touching the owned nodes
constexpr int PF_MAX_ROWS = 128;   // owned (node, element) rows
constexpr int PF_MAX_NODES = 64;   // owned nodes
constexpr int PF_MAX_DEG = 8;      // elements per node
constexpr int PF_MAX_CONTRIB = 4;  // elements sharing one off-diagonal node pair
constexpr int PF_MAX_OUT = 4096;   // CSR values of the owned rows (sum of 13 nb + 1)
constexpr int PF_MAX_BLOCKS = 512; // owned node-pair blocks
struct PlateMaps {
  int n_patches = 0;
  // per patch: {elem_off, node_off, row_off, blk_off}, {n_elems, n_owned, n_rows, n_blocks},
  //            {n_values, 0, 0, 0}
  std::vector<uint32_t> hdr;             // 12 per patch
  std::vector<uint32_t> pconn;           // 4 per patch element: node indices
  std::vector<uint32_t> peid;            // per patch element: element id (insertion order within the patch)
  std::vector<uint32_t> rows;            // per owned row: patch-local element | p << 8
  std::vector<int64_t> own_base;         // per owned node: CSR position of its first row
  std::vector<uint32_t> own_info;        // deg | nb << 8 | kself << 16 | first row << 24
  std::vector<uint32_t> own_boff;        // first block of the owned node within the patch
  std::vector<uint32_t> own_ooff;        // first CSR value of the owned node within the patch
  std::vector<uint32_t> blk;             // per node-pair block: count | (r:3, q:2) << (3 + 5 m) |
                                         //   owned-node slot << 23 | diagonal flag << 29
};
// wg_per_cu: resident workgroups per CU that loop over the patches (0 = one
// workgroup per patch)
std::string upload_plate_maps(Device* d, const PlateMaps& pm, int wg_per_cu);

// Block-lane fused plate path (plate_blocks.hip): same preconditions as the
// maps above; one lane per owned node-pair block.
constexpr int PB_MAX_ELEMS = 46;    // elements touching a patch
constexpr int PB_MAX_BLOCKS = 256;  // owned node-pair blocks of a patch
struct PlateBlockMaps {
  int n_patches = 0;
  std::vector<uint32_t> hdr;    // per patch: elem_off, blk_off, n_elems, n_blocks
  std::vector<uint32_t> pconn;  // 4 per patch element: node indices
  std::vector<uint32_t> peid;   // per patch element: element id (insertion order within the patch)
  // per block, 4 words: CSR position of row 6a of the owned node a;
  // k | nb << 5 | self << 10 | contributors << 11 (k: block within a's row, nb: blocks of a);
  // contributors in element-insertion order: self blocks one byte each
  // (patch element | p << 6, words 3 and 4), others 10 bits each
  // (patch element | p << 6 | q << 8, three in word 3, the fourth in word 4)
  std::vector<uint32_t> blk;
};
std::string upload_plate_block_maps(Device* d, const PlateBlockMaps& pm, int wg_per_cu);
std::string set_plate_block_grid(Device* d, int wg_per_cu);
std::string assemble_plate_blocks(Device* d, double* val, void* stream, bool timed = true);

// Balanced-lane fused plate path (plate_lanes.hip): same preconditions as the
// maps above, and every node in at most four elements with at most nine
// node-pair blocks.  One wave per patch of owned nodes.
constexpr int PL_NODES = 16;         // owned nodes of a wave patch
constexpr int PL_ELEMS = 32;         // elements touching a wave patch
constexpr int PL_TRIPS = 4;          // contributions per lane
constexpr int PL_MAX_NB = 9;         // node-pair blocks per owned node
constexpr int PL_STAGE_VALUES = 1920;  // staged values of a patch: its nodes' 13 nb + 1 values and a parity pad per run
struct PlateLaneMaps {
  int n_waves = 0;
  // per wave PL_ELEMS element slots (insertion order within the wave; unused
  // slots have element id 0xFFFFFFFF): 4 node indices, element id
  std::vector<uint32_t> sconn;
  std::vector<uint32_t> seid;
  // per wave PL_NODES x 2: CSR position of row 6a of owned node s (0 past the
  // last owned node), nb | (n_elems | n_nodes << 6 | trips << 11) << 4 |
  // staging position << 18
  std::vector<uint32_t> nodes;
  // distinct lane-descriptor sets, 64 lanes x 2 words (64 bits) each: up to
  // four items of 13 bits (wave element | p << 5 | q << 7 | block k << 9) in
  // bits 0..51, item count in 52..54, owned-node slot in 55..58, nb in 59..62.
  // Items are grouped by block, insertion order within.  Waves with the same
  // set (all interior patches of a structured mesh) share one copy.
  std::vector<uint32_t> lanes;
  std::vector<uint32_t> lane_set;  // per wave: its set in `lanes`
};
std::string upload_plate_lane_maps(Device* d, const PlateLaneMaps& pm);
// E, nu and k shared by every plate (uniform) or per element: sets the slot
// layout (8 or 10 doubles) and marks the slots for refilling
std::string set_plate_lane_material(Device* d, bool uniform, const double mat[3]);
std::string assemble_plate_lanes(Device* d, double* val, void* stream, bool timed = true);

// Corner-lane fused plate path (plate_corners.hip): same preconditions as the
// maps above, every node in at most four elements, every off-diagonal node
// pair in at most two, and each element corner sharing at most two of its
// blocks.  Uses the same device buffers as the block-lane path.
constexpr int PC_NODES = 64;       // owned nodes of a patch (four lanes each)
constexpr int PC_MAX_ELEMS = 64;   // elements touching a patch
struct PlateCornerMaps {
  int n_patches = 0;
  std::vector<uint32_t> hdr;    // per patch: elem_off, n_elems
  std::vector<uint32_t> pconn;  // 4 per patch element: node indices
  std::vector<uint32_t> peid;   // per patch element: element id (insertion order within the patch)
  // per patch 4 PC_NODES lanes (node slot n: lanes 4 n .. 4 n + 3, the node's
  // elements in insertion order), 4 words each:
  //   CSR position of row 6a of the owned node a;
  //   patch element (127: idle lane) | p << 7 | q1 << 9 | q2 << 11 | q3 << 13 | nb << 15 |
  //     deg << 20 (lane 0 only) | lane 0 << 23 | store slot 1 << 24 | pull into slot 2 << 25 |
  //     source lane in the quad << 26;
  //   k0 | k1 << 5 | k2 << 10 | k3 << 15 (block of each slot within a's rows);
  //   0
  // Slot 0 is the corner itself (q0 = p); slot 1 is the block this lane
  // exports (if any), slot 2 the block it completes with the pulled one.
  std::vector<uint32_t> lane;
};
std::string upload_plate_corner_maps(Device* d, const PlateCornerMaps& pm, int wg_per_cu);
std::string assemble_plate_corners(Device* d, double* val, void* stream, bool timed = true);

// Bar owner path (bar_owner.hip): meshes of trusses and beams in which every
// off-diagonal node pair belongs to one element.  One thread per (node, row).
struct BarMaps {
  std::vector<uint32_t> bid[2];   // trusses, beams: per element node pair (2 p + q): global block
  std::vector<uint64_t> bmask;    // per block: union structural mask
  std::vector<uint16_t> boff;     // per block and row: first entry of the block within the row
  std::vector<uint32_t> inc_ptr;  // N + 1
  std::vector<uint32_t> inc;      // per node, insertion order: type << 31 | element << 1 | p
  // per workgroup: first node, node count.  The first n_bar_ranges go to the
  // bar kernel (<= 40 nodes, <= 256 entries, <= 5120 values), the rest, whose
  // nodes touch trusses only, to the truss kernel (<= 85 nodes, <= 512 entries)
  std::vector<uint32_t> ranges;
  size_t n_bar_ranges = 0;
  // bar-kernel entries in range order: slot s holds the staged constants of
  // entry slot_inc[s] (an index into inc); range r owns slots slot_ptr[r] ..
  std::vector<uint32_t> slot_inc, slot_ptr;
};
constexpr int BAR_WG_NODES = 40, BAR_WG_ENTRIES = 256, BAR_WG_VALUES = 5120;
constexpr int TRUSS_WG_NODES = 85, TRUSS_WG_ENTRIES = 512;
std::string upload_bar_maps(Device* d, const BarMaps& bm);
std::string assemble_bar_owner(Device* d, double* val, void* stream, bool timed = true);

// Value kernels.  `val` is nnz doubles of device memory, `stream` a hipStream_t.
// `timed` records the model's events around the kernels (off inside graph capture).
std::string assemble_generic(Device* d, double* val, void* stream, bool timed = true);
std::string assemble_plate_fused(Device* d, double* val, void* stream, bool timed = true);
double last_ms(Device* d);
// launch shapes (the autotuner's knobs): resident workgroups per CU of the fused
// plate kernel (0 = one per patch); block size of the generic kernels
std::string set_plate_grid(Device* d, int wg_per_cu);
void set_generic_block(Device* d, int threads);

// Batched assembly helpers (batch.hip)
struct Graph;
struct Timer;
std::string stream_create(void*& s);
void stream_destroy(void* s);
std::string stream_sync(void* s);
std::string event_create(void*& ev);
void event_destroy(void* ev);
// `waiter` waits for the work enqueued on `signaler` so far (through `ev`)
std::string stream_wait(void* waiter, void* signaler, void* ev);
std::string graph_begin(void* stream);
std::string graph_end(void* stream, Graph*& g);
std::string graph_launch(Graph* g, void* stream);
void graph_destroy(Graph* g);
Timer* timer_create();
void timer_destroy(Timer* t);
std::string timer_start(Timer* t, void* stream);
std::string timer_stop(Timer* t, void* stream, double& ms);  // synchronizes
std::string timer_end(Timer* t, void* stream);  // records only
double last_ms_between(Timer* t);                // synchronizes on the end event
std::string copy_to_host(void* dst, const void* src_device, size_t bytes);
// architecture name of the current device ("gfx950" without feature suffixes)
std::string device_arch();
int64_t* row_ptr_mut(Device* d);

// Adjoint maps (adjoint.hip)
struct AdjointMaps {
  std::vector<uint32_t> bid[3];   // per element node pair (p * nodes_per + q): global node-pair block
  std::vector<uint64_t> bmask;    // per block: union structural mask (bit 6 i + j)
  std::vector<uint16_t> boff;     // per block and row i: first entry of the block within the CSR row
  std::vector<uint32_t> inc_ptr;  // N + 1
  // per (element, node) gradient row: its slot in node-major order (the rows
  // of node a fill slots inc_ptr[a] .. inc_ptr[a + 1] - 1 in element-insertion
  // order), so the kernels write each row where the node sum reads it
  std::vector<uint32_t> xpos;
  int64_t xoff[3] = {0, 0, 0};    // first gradient row of each element type (beams have none)
  int64_t n_xs_rows = 0;
  bool flat = false;              // plates only, identity frames, standard flat block masks
  // flat: per (element, p) two words: the CSR position of row 6 a of corner
  // p's node a; the block index within a's rows of each q (6 bits each) | nb << 24
  std::vector<uint32_t> kpack;
  // flat: the elements in the order the flat backward takes them, 16 per
  // workgroup (Morton order of node 3: compact tiles, whose lanes' gathers
  // share cache lines)
  std::vector<uint32_t> eorder;
};
std::string upload_adjoint_maps(Device* d, const AdjointMaps& am);
// mode 0: g = dL/dK (nnz values, CSR order); mode 1: g = nodal displacements
// (6 N) and the outputs are -u_e^T (dK_e/dp) u_e.  dparam[t]: n_t x params_per
// (or null); dxyz: N x 3 (or null; trusses and plates only).
// wide: dL/dK block gathers with 16-byte range-checked loads when the values span under 4 GiB
std::string run_adjoint(Device* d, int mode, const double* g, double* const dparam[3], double* dxyz, void* stream,
                        bool wide = true);
double last_backward_ms(Device* d);

// Device-to-device replacement of one element type's parameters or of the node
// coordinates (then synchronous); the host mirrors are refreshed from the
// device.  The caller reruns run_setup for the frames.
std::string replace_params(Device* d, int type, const double* src, double* host_copy, void* stream);
std::string replace_nodes(Device* d, const double* src, double* host_copy, void* stream);

double* values(Device* d);
const int64_t* row_ptr(Device* d);
const int32_t* col_idx(Device* d);
std::string copy_csr_to_host(Device* d, int64_t n_rows, int64_t nnz, int64_t* rp, int32_t* ci, double* v);

// Sparse separation (separate.hip).  upload_bc: per global dof the
// constrained flag, imposed displacement and force.  separate_sparse
// partitions the current CSR values; err_row is the first constrained dof
// without stiffness (-1 if none).
struct SepResult {
  int64_t err_row = -1, n_aa = 0, n_bb = 0;
  int64_t nnz[4] = {0, 0, 0, 0};  // K_aa, K_ab, K_ba, K_bb
};
std::string upload_bc(Device* d, const uint8_t* constrained, const double* disp, const double* force, size_t n);
std::string separate_sparse(Device* d, SepResult& res, void* stream, bool timed, int lanes = 0);
std::string separated_csr(Device* d, int which, const int64_t** rp, const int32_t** ci, const double** v);
std::string separated_vectors(Device* d, const int64_t** aa, const int64_t** bb, const double** r_a,
                              const double** u_b);
void separation_shape(Device* d, int64_t rows[4], int64_t nnz[4]);
void invalidate_separation(Device* d);
double last_separate_ms(Device* d);

// Sparse global analysis on the last separation (solve.hip): b = r_a - K_ab
// u_b, K_aa u_a = b by PCG (precond 0: Jacobi, 1: node-block Jacobi),
// r_r = K_ba u_a + K_bb u_b - r_b.  Synchronous.
struct SolveResult {
  int64_t iterations = 0;
  bool converged = false;
  double residual = 0.0, stop = 0.0;
};
// wide: the update loop's SpMV reads four values per lane and step with 16-byte range-checked loads
// when K_aa's values span under 4 GiB
std::string solve_pcg(Device* d, int precond, int max_iter, double rel_tol, double abs_tol, void* stream,
                      SolveResult& res, bool wide = true);
// The same solve with every work buffer in one static device array of
// STATIC_SCRATCH_BYTES per GPU (no allocation); u_a (n_aa) and r_r (n_bb) are
// caller device memory.  The model's stored solution is not changed.
constexpr size_t STATIC_SCRATCH_BYTES = size_t(64) << 20;
std::string solve_pcg_static(Device* d, int precond, int max_iter, double rel_tol, double abs_tol, void* stream,
                             double* u_a, double* r_r, SolveResult& res, bool wide = true);
std::string solution(Device* d, const double** u_a, const double** r_r, int64_t* n_aa, int64_t* n_bb);
double last_solve_ms(Device* d);

// Element post-processing (forces.hip): global displacements of the last
// solve (imposed values with u_a at the k_aa indexes, device-owned) and the
// element forces from any device displacement vector (6 per node); out[t]
// holds 1 (truss), 10 (beam) or 8 (plate) values per element, or is null.
std::string solution_displacements(Device* d, void* stream, const double** u);
std::string element_forces(Device* d, const double* u, double* const out[3], void* stream);
// Nodal loads of uniformly distributed loads (type 1: beams, 2 per item;
// type 2: plates, 4 per item) on the elements at positions pos; host arrays in
// and out, synchronous.
std::string nodal_loads(Device* d, int type, const uint32_t* pos, const double* values, size_t n, double* out);

}  // namespace femdev
