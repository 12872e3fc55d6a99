// fem_host.cpp -- host side of the assembly library: the crate's FEM<V> model
// (nodes, elements, validation in the crate's order and with its messages),
// the sparsity-pattern / element-to-slot map builder, and the C ABI of
// include/fem_amd.h.  All element arithmetic runs on the device
// (fem_device.hip, plate_fused.hip); the host only validates and indexes.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <array>
#include <map>
#include <charconv>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fem_amd.h"
#include "element_math.hpp"
#include "fem_device.hpp"

using femdev::BEAM;
using femdev::PLATE;
using femdev::TRUSS;

namespace {

// Rust `{:?}` of an f64 (shortest round-trip digits; exponential form when
// 0 < |x| < 1e-4 or |x| >= 1e16), used in the crate's error messages.
std::string rust_debug(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-inf" : "inf";
  char buf[64];
  auto r = std::to_chars(buf, buf + 64, std::fabs(x), std::chars_format::scientific);
  std::string s(buf, r.ptr);
  size_t e = s.find('e');
  int ex = std::stoi(s.substr(e + 1));
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = std::signbit(x) ? "-" : "";
  double ax = std::fabs(x);
  if (ax == 0.0) return out + "0.0";
  if (ax < 1e-4 || ax >= 1e16) {
    out += digits[0];
    if (digits.size() > 1) out += "." + digits.substr(1);
    return out + "e" + std::to_string(ex);
  }
  int point = ex + 1;
  if (point <= 0) return out + "0." + std::string(size_t(-point), '0') + digits;
  if (int(digits.size()) <= point) return out + digits + std::string(size_t(point) - digits.size(), '0') + ".0";
  return out + digits.substr(0, size_t(point)) + "." + digits.substr(size_t(point));
}

std::string u(uint64_t v) { return std::to_string(v); }

uint64_t coord_key(double x, double y, double z) {
  auto h = [](double v) {
    if (v == 0.0) v = 0.0;
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
  };
  return h(x) * 0x9E3779B97F4A7C15ull ^ (h(y) + 0x7F4A7C15ull) * 0xC2B2AE3D27D4EB4Full ^ h(z);
}
uint64_t pair_key(uint32_t a, uint32_t b) { return a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a); }

struct ElementSet {
  std::vector<uint32_t> number;
  std::vector<uint32_t> conn;    // node indices
  std::vector<double> params;
  std::vector<uint64_t> seq;     // global insertion sequence
  std::unordered_map<uint32_t, uint32_t> by_number;
  size_t size() const { return number.size(); }
};

}  // namespace

struct fem_model {
  double rel_tol = 0, abs_tol = 0;
  uint32_t nodes_number = 0;
  std::string err;
  std::unordered_map<uint32_t, uint32_t> node_idx;
  std::vector<uint32_t> node_number;
  std::vector<double> xyz;
  std::unordered_map<uint64_t, std::vector<uint32_t>> coord_hash;
  ElementSet set[3];
  std::unordered_map<uint64_t, uint32_t> truss_pairs, beam_pairs;
  std::vector<std::vector<uint32_t>> node_plates;
  uint64_t seq = 0;
  // device state
  femdev::Device* dev = nullptr;
  bool dirty = true;           // host model changed since the last device upload
  bool pattern_valid = false;
  std::vector<uint16_t> rzero[3];
  int64_t nnz = 0;
  std::string path = "none";
  int pf_nodes = 64;   // fused plate: owned nodes per patch (tuning option)
  int pf_wg_per_cu = -1;  // fused plate: resident workgroups per CU looping over patches (0 = one per
                          // patch, -1 = the kernel's default: 2 for plate_fused, 0 for plate_blocks)
  int pf_kernel = 3;      // fused plate kernel: 0 = row staging (plate_fused), 1 = one lane per block
                          // (plate_blocks), 2 = one lane per element corner (plate_corners),
                          // 3 = balanced lanes (plate_lanes; falls back to plate_blocks)
  bool force_generic = false;
  bool wide_loads = true;  // backward block gathers and PCG SpMV: 16-byte loads when the values fit 32-bit offsets
  bool time_assembly = true;  // events around each assembly launch (fem_last_assembly_ms)
  int pl_order = 0;  // balanced-lane plate patches: 0 = automatic, 1 = node-index order, 2 = Morton order
  // node-pair blocks of the pattern (kept for the adjoint maps): node a owns
  // blocks pat_bptr[a] .. pat_bptr[a + 1], sorted by neighbour pat_bnode[k]
  std::vector<uint64_t> pat_bptr;
  std::vector<uint32_t> pat_bnode;
  std::vector<uint64_t> pat_bmask;
  std::vector<uint16_t> pat_boff;
  bool adj_valid = false;
  std::vector<int64_t> row_ptr_host;  // host copy of the pattern's row_ptr
  std::vector<uint32_t> node_part;    // batch groups: mesh of each node (major key of the patch order)
  // boundary conditions per global dof (methods_for_bc_data_handle.rs): the
  // crate's imposed_constraints, displacements_vector and forces_vector
  std::vector<uint8_t> bc_cons;
  std::vector<double> bc_disp, bc_force;
  bool bc_dirty = true;
  bool own_vals = false;  // the library's own CSR values hold the current stiffness
  int sep_lanes = 0;      // separation: lanes per CSR row (tuning option, 0 = automatic)
  // forward + backward: the backward runs on a second stream, forked from and
  // joined to the caller's stream
  void* side_stream = nullptr;
  void* fork_ev = nullptr;
  void* join_ev = nullptr;
};

// A batch of independent meshes assembled together (no counterpart in the
// crate).  Meshes are copied into per-group disjoint-union models: flat
// plate-only meshes (fused plate kernel) and all others (generic path).  A
// mesh's nodes are a contiguous index range of its group, so its CSR rows and
// values are contiguous slices of the group CSR and its values are summed in
// its own element-insertion order, exactly as when it is assembled alone.
struct fem_batch {
  struct Mesh {
    int group = -1;
    size_t node_off = 0, n_nodes = 0;
    int64_t val_off = 0, nnz = 0;
  };
  std::string err;
  std::vector<fem_model*> groups;
  std::vector<Mesh> meshes;
  void* stream = nullptr;
  // group k > 0 is captured on side[k - 1], forked from and joined to `stream`
  // by events, so the graph runs the groups as parallel branches
  std::vector<void*> side, fork_ev, join_ev;
  femdev::Graph* graph = nullptr;
  femdev::Timer* timer = nullptr;
  bool ready = false, timed = false;
  std::string config = "untuned";
  // 0: every group on its own graph branch (concurrent); 1: the groups one
  // after another on one stream
  int schedule = 0;
};

namespace {

int fail(fem_model* m, int code, const std::string& msg) {
  m->err = msg;
  return code;
}

void clear_model(fem_model* m, uint32_t nodes_number) {
  m->nodes_number = nodes_number;
  m->node_idx.clear();
  m->node_number.clear();
  m->xyz.clear();
  m->coord_hash.clear();
  for (auto& s : m->set) s = ElementSet();
  m->truss_pairs.clear();
  m->beam_pairs.clear();
  m->node_plates.clear();
  m->seq = 0;
  m->dirty = true;
  m->pattern_valid = false;
  m->adj_valid = false;
  m->nnz = 0;
  for (auto& r : m->rzero) r.clear();
  m->bc_cons.clear();
  m->bc_disp.clear();
  m->bc_force.clear();
  m->bc_dirty = true;
  m->own_vals = false;
  if (m->dev) femdev::invalidate_separation(m->dev);
}

bool node_exists(fem_model* m, uint32_t number) {  // methods_for_node_data_handle.rs:81-86
  if (!m->node_idx.count(number)) {
    m->err = "Node with number " + u(number) + " does not exist!";
    return false;
  }
  return true;
}

const double* node_xyz(const fem_model* m, uint32_t idx) { return &m->xyz[3 * size_t(idx)]; }

bool positive_or(fem_model* m, double v, const char* what, double shown) {
  if (v <= 0.0 || std::isnan(v)) {
    if (!(v <= 0.0)) return true;  // NaN compares false in the crate too
    m->err = std::string(what) + " " + rust_debug(shown) + " is less or equal to zero!";
    return false;
  }
  return true;
}

// -------------------------------------------------------- plate geometry checks
// quadrilateral_4n_element_functions.rs:14-86
bool quad_on_line(const double* p[4]) {
  const int pairs[4][3] = {{0, 1, 3}, {1, 0, 2}, {2, 1, 3}, {3, 2, 0}};
  for (auto& pr : pairs) {
    const double *o = p[pr[0]], *a = p[pr[1]], *b = p[pr[2]];
    double v1[3] = {a[0] - o[0], a[1] - o[1], a[2] - o[2]};
    double v2[3] = {b[0] - o[0], b[1] - o[1], b[2] - o[2]};
    double c[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
    if (fem::norm3(c[0], c[1], c[2]) == 0.0) return true;
  }
  return false;
}
// quadrilateral_4n_element_functions.rs:88-129
bool quad_on_plane(const double* p[4], double abs_tol) {
  const double *p1 = p[0], *p2 = p[1], *p3 = p[2], *p4 = p[3];
  double v32[3] = {p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
  double v34[3] = {p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
  double a = v32[1] * v34[2] - v32[2] * v34[1], b = v32[2] * v34[0] - v32[0] * v34[2],
         c = v32[0] * v34[1] - v32[1] * v34[0];
  double d = -1.0 * (a * p3[0] + b * p3[1] + c * p3[2]);
  return fem::cwt(a * p1[0] + b * p1[1] + c * p1[2] + d, abs_tol) == 0.0;
}

// convex_hull_on_plane.rs:14-214: polar-angle quicksort about the lowest point,
// then removal of non-left turns; the crate requires 4 hull points.
struct HP {
  uint32_t n;
  double x, y;
};
double hp_angle(const HP& p) {
  double rhs = std::sqrt((0.0 - p.x) * (0.0 - p.x) + (0.0 - p.y) * (0.0 - p.y));
  double sp = (0.0 - 1.0) * (0.0 - p.x) + (0.0 - 0.0) * (0.0 - p.y);
  return std::acos(sp / (1.0 * rhs)) * (180.0 / M_PI);
}
void hp_sort(HP* a, long low, long high) {
  if (low >= high) return;
  long pivot = high, store = low - 1, last = high;
  while (true) {
    store += 1;
    while (hp_angle(a[store]) < hp_angle(a[pivot])) store += 1;
    last -= 1;
    while (last >= 0 && hp_angle(a[last]) > hp_angle(a[pivot])) last -= 1;
    if (store >= last) break;
    std::swap(a[store], a[last]);
  }
  std::swap(a[store], a[pivot]);
  hp_sort(a, low, store - 1);
  hp_sort(a, store + 1, high);
}
std::vector<HP> hull_points(std::vector<HP> u) {
  double sx = u[0].x, my = u[0].y;
  size_t mp = 0;
  for (size_t i = 0; i < u.size(); ++i)
    if (u[i].y < my) { sx = u[i].x; my = u[i].y; mp = i; }
  std::swap(u[0], u[mp]);
  for (auto& p : u) { p.x -= sx; p.y -= my; }
  hp_sort(u.data() + 1, 0, long(u.size()) - 2);
  size_t i = 0;
  while (i + 2 < u.size()) {
    double area = (u[i + 1].x - u[i].x) * (u[i + 2].y - u[i].y) - (u[i + 1].y - u[i].y) * (u[i + 2].x - u[i].x);
    if (area <= 0.0) u.erase(u.begin() + long(i) + 1);
    else i += 1;
  }
  return u;
}
// quadrilateral_4n_element_functions.rs:173-250: the hull's point numbers in
// the crate's order; false when the frame of nodes 2-4 does not exist
bool quad_hull(const uint32_t num[4], const double* p[4], double abs_tol, std::vector<uint32_t>& out) {
  double R[9];
  if (!fem::quad_rotation(p[1], p[2], p[3], abs_tol, R)) return false;
  auto tr = [&](const double* q, double& x, double& y) {
    double d[3] = {q[0] - p[2][0], q[1] - p[2][1], q[2] - p[2][2]};
    x = (R[0] * d[0] + R[1] * d[1]) + R[2] * d[2];
    y = (R[3] * d[0] + R[4] * d[1]) + R[5] * d[2];
  };
  std::vector<HP> hp(4);
  hp[0].n = num[0]; tr(p[0], hp[0].x, hp[0].y);
  hp[1].n = num[1]; tr(p[1], hp[1].x, hp[1].y);
  hp[2] = {num[2], 0.0, 0.0};
  hp[3].n = num[3]; tr(p[3], hp[3].x, hp[3].y);
  out.clear();
  for (const HP& q : hull_points(hp)) out.push_back(q.n);
  return true;
}
// number of hull points, -1 when no frame exists
int quad_hull_points(const uint32_t num[4], const double* p[4], double abs_tol) {
  std::vector<uint32_t> h;
  if (!quad_hull(num, p, abs_tol, h)) return -1;
  return int(h.size());
}

// -------------------------------------------------------- structural patterns
// Node-pair block masks: bit 6*i+j of the 6x6 block of K_e for local nodes (p, q).
using Mask = uint64_t;
inline bool bit(Mask m, int i, int j) { return (m >> (6 * i + j)) & 1ull; }
inline void setb(Mask& m, int i, int j) { m |= 1ull << (6 * i + j); }

// structural pattern of R6^T L R6 from the zero pattern of R
Mask rotate_mask(Mask L, uint16_t rz) {
  auto r = [&](int k, int i) { return (rz >> (3 * k + i)) & 1; };  // R[k][i] != 0
  Mask out = 0;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          bool nz = false;
          for (int k = 0; k < 3 && !nz; ++k)
            for (int l = 0; l < 3 && !nz; ++l)
              nz = r(k, i) && bit(L, 3 * bi + k, 3 * bj + l) && r(l, j);
          if (nz) setb(out, 3 * bi + i, 3 * bj + j);
        }
  return out;
}
Mask element_pair_mask(int type, uint16_t rz, int p, int q) {
  Mask L = 0;
  if (type == TRUSS) {
    // truss.rs: only u-u is non-zero locally; global block R0^T R0 scaled
    Mask out = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (((rz >> i) & 1) && ((rz >> j) & 1)) setb(out, i, j);
    return out;
  }
  if (type == BEAM) {
    // beam.rs:495-563: u | (v, tw) | (w, tv) | tu | tv | tw
    setb(L, 0, 0);
    for (int a : {1, 5})
      for (int b : {1, 5}) setb(L, a, b);
    for (int a : {2, 4})
      for (int b : {2, 4}) setb(L, a, b);
    setb(L, 3, 3);
    Mask out = rotate_mask(L, rz);
    // shear isotropy: kv (r1 r1^T + r2 r2^T) = kv (I - r0 r0^T) in the translational block
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (i != j) {
          bool nz = ((rz >> i) & 1) && ((rz >> j) & 1);
          out &= ~(1ull << (6 * i + j));
          if (nz) setb(out, i, j);
        }
    return out;
  }
  // plate.rs:160-511: membrane (u, v), bending + MITC4 shear (w, tx, ty), drilling on the diagonal
  for (int a : {0, 1})
    for (int b : {0, 1}) setb(L, a, b);
  for (int a : {2, 3, 4})
    for (int b : {2, 3, 4}) setb(L, a, b);
  if (p == q) setb(L, 5, 5);
  return rotate_mask(L, rz);
}

// -------------------------------------------------------- pattern builder
struct PairRef {
  uint32_t b;
  Mask mask;
};

// -------------------------------------------------------- plate fused maps
constexpr Mask plate_off_mask() {
  Mask mk = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) mk |= 1ull << (6 * a + b);
  for (int a = 2; a < 5; ++a)
    for (int b = 2; b < 5; ++b) mk |= 1ull << (6 * a + b);
  return mk;
}

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  uint64_t x = v & 0x1FFFFFull;
  x = (x | x << 32) & 0x1F00000000FFFFull;
  x = (x | x << 16) & 0x1F0000FF0000FFull;
  x = (x | x << 8) & 0x100F00F00F00F00Full;
  x = (x | x << 4) & 0x10C30C30C30C30C3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Maps of the fused plate kernel: only plates, every frame exactly the
// identity, standard flat-plate block masks, bounded valence.  Returns false
// (and the generic path is used) otherwise.
bool build_plate_maps(fem_model* m, const std::vector<uint64_t>& bptr, const std::vector<PairRef>& blocks,
                      const std::vector<int64_t>& row_ptr, femdev::PlateMaps& pm) {
  const auto& S = m->set[PLATE];
  if (m->set[TRUSS].size() || m->set[BEAM].size() || S.size() == 0 || S.size() >= (1u << 30)) return false;
  if (row_ptr.back() * 8 >= (int64_t(1) << 32)) return false;  // 32-bit byte offsets of the buffer stores
  for (uint16_t rz : m->rzero[PLATE])
    if (!(rz & 0x4000)) return false;
  const size_t N = m->nodes_number, E = S.size();
  const Mask MOFF = plate_off_mask(), MSELF = plate_off_mask() | (1ull << 35);
  // node -> (element | p << 30) in insertion order
  std::vector<uint32_t> deg(N + 1, 0);
  for (size_t e = 0; e < E; ++e)
    for (int p = 0; p < 4; ++p) deg[S.conn[4 * e + p]]++;
  std::vector<uint32_t> nptr(N + 1, 0);
  for (size_t a = 0; a < N; ++a) {
    if (deg[a] > uint32_t(femdev::PF_MAX_DEG)) return false;
    nptr[a + 1] = nptr[a] + deg[a];
  }
  std::vector<uint32_t> nel(nptr[N]);
  {
    std::vector<uint32_t> f(nptr.begin(), nptr.end() - 1);
    for (size_t e = 0; e < E; ++e)
      for (int p = 0; p < 4; ++p) nel[f[S.conn[4 * e + p]]++] = uint32_t(e) | uint32_t(p) << 30;
  }
  // Morton order of the nodes that carry elements
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t a = 0; a < N; ++a)
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], m->xyz[3 * a + k]); hi[k] = std::max(hi[k], m->xyz[3 * a + k]); }
  std::vector<std::pair<uint64_t, uint32_t>> order;
  order.reserve(N);
  for (size_t a = 0; a < N; ++a) {
    if (!deg[a]) continue;
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double span = hi[k] - lo[k];
      uint64_t qv = span > 0 ? uint64_t((m->xyz[3 * a + k] - lo[k]) / span * 2097151.0) : 0;
      code |= spread3(qv) << k;
    }
    order.push_back({code, uint32_t(a)});
  }
  if (m->node_part.size() == N) {
    // batch groups: meshes one after another, Morton order within each
    std::stable_sort(order.begin(), order.end(), [&](const std::pair<uint64_t, uint32_t>& x,
                                                     const std::pair<uint64_t, uint32_t>& y) {
      const uint32_t px = m->node_part[x.second], py = m->node_part[y.second];
      return px != py ? px < py : x < y;
    });
  } else {
    std::sort(order.begin(), order.end());
  }
  // greedy patches over the Morton order, within the kernel's limits
  const uint32_t max_nodes = uint32_t(std::min(m->pf_nodes, femdev::PF_MAX_NODES));
  std::vector<uint32_t> cur, cur_elems;
  uint32_t cur_rows = 0, cur_blocks = 0, cur_out = 0;
  std::vector<int32_t> elem_local(E, -1);
  auto flush = [&]() -> bool {
    if (cur.empty()) return true;
    std::sort(cur.begin(), cur.end());
    std::sort(cur_elems.begin(), cur_elems.end());  // insertion order
    for (size_t i = 0; i < cur_elems.size(); ++i) elem_local[cur_elems[i]] = int32_t(i);
    const uint32_t elem_off = uint32_t(pm.peid.size()), node_off = uint32_t(pm.own_info.size()),
                   row_off = uint32_t(pm.rows.size()), blk_off = uint32_t(pm.blk.size());
    uint32_t slot = 0, nloc = 0, ooff = 0;
    for (uint32_t a : cur) {
      const uint32_t* el = &nel[nptr[a]];
      uint32_t d = deg[a];
      uint64_t k0 = bptr[a], k1 = bptr[a + 1];
      uint32_t nb = uint32_t(k1 - k0);
      int kself = -1;
      const uint32_t boff = uint32_t(pm.blk.size()) - blk_off;
      for (uint64_t k = k0; k < k1; ++k) {
        uint32_t b = blocks[k].b;
        if (b == a) {
          if (blocks[k].mask != MSELF) return false;
          kself = int(k - k0);
          pm.blk.push_back(nloc << 23 | 1u << 29);
          continue;
        }
        if (blocks[k].mask != MOFF) return false;
        uint32_t desc = 0, cnt = 0;
        for (uint32_t r = 0; r < d; ++r) {
          uint32_t e = el[r] & 0x3FFFFFFFu;
          for (int q = 0; q < 4; ++q)
            if (S.conn[4 * size_t(e) + q] == b) {
              if (cnt == uint32_t(femdev::PF_MAX_CONTRIB)) return false;
              desc |= (r | uint32_t(q) << 3) << (3 + 5 * cnt);
              cnt++;
            }
        }
        if (!cnt) return false;
        pm.blk.push_back(desc | cnt | nloc << 23);
      }
      if (kself < 0) return false;
      if (row_ptr[6 * size_t(a) + 6] - row_ptr[6 * size_t(a)] != int64_t(13 * nb + 1)) return false;
      pm.own_base.push_back(row_ptr[6 * size_t(a)]);
      pm.own_info.push_back(d | nb << 8 | uint32_t(kself) << 16 | slot << 24);
      pm.own_boff.push_back(boff);
      pm.own_ooff.push_back(ooff);
      ooff += 13 * nb + 1;
      for (uint32_t r = 0; r < d; ++r) {
        uint32_t e = el[r] & 0x3FFFFFFFu, p = el[r] >> 30;
        pm.rows.push_back(uint32_t(elem_local[e]) | p << 8);
      }
      slot += d;
      nloc++;
    }
    for (uint32_t e : cur_elems) {
      pm.peid.push_back(e);
      for (int q = 0; q < 4; ++q) pm.pconn.push_back(S.conn[4 * size_t(e) + q]);
      elem_local[e] = -1;
    }
    const uint32_t h[12] = {elem_off, node_off, row_off, blk_off, uint32_t(cur_elems.size()), uint32_t(cur.size()),
                            uint32_t(pm.rows.size()) - row_off, uint32_t(pm.blk.size()) - blk_off, ooff, 0, 0, 0};
    pm.hdr.insert(pm.hdr.end(), h, h + 12);
    pm.n_patches++;
    cur.clear();
    cur_elems.clear();
    cur_rows = cur_blocks = cur_out = 0;
    return true;
  };
  std::vector<uint8_t> in_patch(E, 0);
  for (auto& o : order) {
    uint32_t a = o.second;
    uint32_t nb = uint32_t(bptr[a + 1] - bptr[a]);
    if (nb > 16) return false;
    uint32_t new_elems = 0;
    for (uint32_t r = 0; r < deg[a]; ++r) new_elems += !in_patch[nel[nptr[a] + r] & 0x3FFFFFFFu];
    if (cur.size() == max_nodes || cur_rows + deg[a] > uint32_t(femdev::PF_MAX_ROWS) ||
        cur_elems.size() + new_elems > size_t(femdev::PF_MAX_ELEMS) || cur_blocks + nb > uint32_t(femdev::PF_MAX_BLOCKS) ||
        cur_out + 13 * nb + 1 > uint32_t(femdev::PF_MAX_OUT)) {
      for (uint32_t e : cur_elems) in_patch[e] = 0;
      if (!flush()) return false;
    }
    cur.push_back(a);
    cur_rows += deg[a];
    cur_blocks += nb;
    cur_out += 13 * nb + 1;
    for (uint32_t r = 0; r < deg[a]; ++r) {
      uint32_t e = nel[nptr[a] + r] & 0x3FFFFFFFu;
      if (!in_patch[e]) { in_patch[e] = 1; cur_elems.push_back(e); }
    }
  }
  for (uint32_t e : cur_elems) in_patch[e] = 0;
  return flush();
}

// Maps of the block-lane plate kernel (one lane per owned node-pair block).
// Called only after build_plate_maps accepted the mesh, which checked the
// preconditions (plates only, identity frames, standard block masks, valence
// and row lengths).
// Plate incidence shared by the fused-path map builders: per node its
// elements (element | corner << 30) in insertion order, and the nodes that
// have elements in Morton order of their coordinates (grouped by node_part
// when the batch set one).
void plate_node_incidence(const fem_model* m, std::vector<uint32_t>& deg, std::vector<uint32_t>& nptr,
                          std::vector<uint32_t>& nel, std::vector<std::pair<uint64_t, uint32_t>>& order) {
  const auto& S = m->set[PLATE];
  const size_t N = m->nodes_number, E = S.size();
  deg.assign(N + 1, 0);
  for (size_t e = 0; e < E; ++e)
    for (int p = 0; p < 4; ++p) deg[S.conn[4 * e + p]]++;
  nptr.assign(N + 1, 0);
  for (size_t a = 0; a < N; ++a) nptr[a + 1] = nptr[a] + deg[a];
  nel.assign(nptr[N], 0);
  {
    std::vector<uint32_t> f(nptr.begin(), nptr.end() - 1);
    for (size_t e = 0; e < E; ++e)
      for (int p = 0; p < 4; ++p) nel[f[S.conn[4 * e + p]]++] = uint32_t(e) | uint32_t(p) << 30;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t a = 0; a < N; ++a)
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], m->xyz[3 * a + k]); hi[k] = std::max(hi[k], m->xyz[3 * a + k]); }
  order.clear();
  order.reserve(N);
  for (size_t a = 0; a < N; ++a) {
    if (!deg[a]) continue;
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
      double span = hi[k] - lo[k];
      uint64_t qv = span > 0 ? uint64_t((m->xyz[3 * a + k] - lo[k]) / span * 2097151.0) : 0;
      code |= spread3(qv) << k;
    }
    order.push_back({code, uint32_t(a)});
  }
  if (m->node_part.size() == N) {
    std::stable_sort(order.begin(), order.end(), [&](const std::pair<uint64_t, uint32_t>& x,
                                                     const std::pair<uint64_t, uint32_t>& y) {
      const uint32_t px = m->node_part[x.second], py = m->node_part[y.second];
      return px != py ? px < py : x < y;
    });
  } else {
    std::sort(order.begin(), order.end());
  }
}

bool build_plate_block_maps(fem_model* m, const std::vector<uint64_t>& bptr, const std::vector<PairRef>& blocks,
                            const std::vector<int64_t>& row_ptr, femdev::PlateBlockMaps& pm) {
  const auto& S = m->set[PLATE];
  const size_t E = S.size();
  std::vector<uint32_t> deg, nptr, nel;
  std::vector<std::pair<uint64_t, uint32_t>> order;
  plate_node_incidence(m, deg, nptr, nel, order);
  std::vector<uint32_t> cur, cur_elems;
  uint32_t cur_blocks = 0;
  std::vector<int32_t> elem_local(E, -1);
  std::vector<uint8_t> in_patch(E, 0);
  auto flush = [&]() -> bool {
    if (cur.empty()) return true;
    std::sort(cur.begin(), cur.end());
    std::sort(cur_elems.begin(), cur_elems.end());  // insertion order
    for (size_t i = 0; i < cur_elems.size(); ++i) elem_local[cur_elems[i]] = int32_t(i);
    const uint32_t elem_off = uint32_t(pm.peid.size()), blk_off = uint32_t(pm.blk.size() / 4);
    for (uint32_t a : cur) {
      const uint32_t* el = &nel[nptr[a]];
      const uint32_t d = deg[a];
      const uint64_t k0 = bptr[a], k1 = bptr[a + 1];
      const uint32_t nb = uint32_t(k1 - k0);
      for (uint64_t k = k0; k < k1; ++k) {
        const uint32_t b = blocks[k].b;
        uint32_t w1 = uint32_t(k - k0) | nb << 5, w2 = 0, w3 = 0;
        if (b == a) {
          w1 |= 1u << 10 | d << 11;
          for (uint32_t r = 0; r < d; ++r) {
            const uint32_t e = el[r] & 0x3FFFFFFFu, p = el[r] >> 30;
            const uint32_t cb = uint32_t(elem_local[e]) | p << 6;
            if (r < 4) w2 |= cb << (8 * r);
            else w3 |= cb << (8 * (r - 4));
          }
        } else {
          uint32_t cnt = 0;
          for (uint32_t r = 0; r < d; ++r) {
            const uint32_t e = el[r] & 0x3FFFFFFFu, p = el[r] >> 30;
            for (uint32_t q = 0; q < 4; ++q)
              if (S.conn[4 * size_t(e) + q] == b) {
                if (cnt == uint32_t(femdev::PF_MAX_CONTRIB)) return false;
                const uint32_t cb = uint32_t(elem_local[e]) | p << 6 | q << 8;
                if (cnt < 3) w2 |= cb << (10 * cnt);
                else w3 = cb;
                cnt++;
              }
          }
          if (!cnt) return false;
          w1 |= cnt << 11;
        }
        pm.blk.insert(pm.blk.end(), {uint32_t(row_ptr[6 * size_t(a)]), w1, w2, w3});
      }
    }
    for (uint32_t e : cur_elems) {
      pm.peid.push_back(e);
      for (int q = 0; q < 4; ++q) pm.pconn.push_back(S.conn[4 * size_t(e) + q]);
      elem_local[e] = -1;
    }
    pm.hdr.insert(pm.hdr.end(), {elem_off, blk_off, uint32_t(cur_elems.size()), uint32_t(pm.blk.size() / 4) - blk_off});
    pm.n_patches++;
    cur.clear();
    cur_elems.clear();
    cur_blocks = 0;
    return true;
  };
  for (auto& o : order) {
    const uint32_t a = o.second;
    const uint32_t nb = uint32_t(bptr[a + 1] - bptr[a]);
    uint32_t new_elems = 0;
    for (uint32_t r = 0; r < deg[a]; ++r) new_elems += !in_patch[nel[nptr[a] + r] & 0x3FFFFFFFu];
    if (cur_blocks + nb > uint32_t(femdev::PB_MAX_BLOCKS) ||
        cur_elems.size() + new_elems > size_t(femdev::PB_MAX_ELEMS)) {
      for (uint32_t e : cur_elems) in_patch[e] = 0;
      if (!flush()) return false;
    }
    cur.push_back(a);
    cur_blocks += nb;
    for (uint32_t r = 0; r < deg[a]; ++r) {
      const uint32_t e = nel[nptr[a] + r] & 0x3FFFFFFFu;
      if (!in_patch[e]) { in_patch[e] = 1; cur_elems.push_back(e); }
    }
  }
  for (uint32_t e : cur_elems) in_patch[e] = 0;
  return flush();
}

// true when every plate has the same E, nu and k, bit for bit (into mat); the
// balanced-lane slots then hold t only
bool plate_material_uniform(const fem_model* m, double mat[3]) {
  const auto& S = m->set[PLATE];
  if (!S.size()) return false;
  const double* P = S.params.data();
  const double first[3] = {P[0], P[1], P[3]};
  for (size_t e = 0; e < S.size(); ++e) {
    const double v[3] = {P[4 * e], P[4 * e + 1], P[4 * e + 3]};
    if (std::memcmp(v, first, sizeof(v)) != 0) return false;
  }
  for (int k = 0; k < 3; ++k) mat[k] = first[k];
  return true;
}

// Maps of the balanced-lane plate kernel (plate_lanes.hip).  Called only after
// build_plate_maps accepted the mesh.  Returns false when a node has more than
// PL_TRIPS elements or more than PL_MAX_NB node-pair blocks (the caller then
// uses the block-lane maps).
bool build_plate_lane_maps(fem_model* m, const std::vector<uint64_t>& bptr, const std::vector<PairRef>& blocks,
                           const std::vector<int64_t>& row_ptr, femdev::PlateLaneMaps& pm) {
  const auto& S = m->set[PLATE];
  const size_t E = S.size();
  if (row_ptr.back() >= (int64_t(1) << 31)) return false;
  std::vector<uint32_t> deg, nptr, nel;
  std::vector<std::pair<uint64_t, uint32_t>> order;
  plate_node_incidence(m, deg, nptr, nel, order);
  struct Item { uint32_t e, p, q, k; };
  // lanes of one node: up to PL_TRIPS items each, blocks never split
  auto node_lanes = [&](uint32_t a, std::vector<std::vector<Item>>& lanes) -> bool {
    lanes.clear();
    const uint32_t d = deg[a];
    const uint64_t k0 = bptr[a], k1 = bptr[a + 1];
    if (d > uint32_t(femdev::PL_TRIPS) || k1 - k0 > uint64_t(femdev::PL_MAX_NB)) return false;
    std::vector<std::vector<Item>> blk(size_t(k1 - k0));
    const uint32_t* el = &nel[nptr[a]];
    for (uint64_t k = k0; k < k1; ++k) {
      const uint32_t b = blocks[k].b;
      for (uint32_t r = 0; r < d; ++r) {  // element insertion order
        const uint32_t e = el[r] & 0x3FFFFFFFu, p = el[r] >> 30;
        for (uint32_t q = 0; q < 4; ++q)
          if (S.conn[4 * size_t(e) + q] == b) blk[size_t(k - k0)].push_back({e, p, q, uint32_t(k - k0)});
      }
      if (blk[size_t(k - k0)].empty() || blk[size_t(k - k0)].size() > size_t(femdev::PL_TRIPS)) return false;
    }
    // first fit, blocks by decreasing contribution count (stable in k)
    std::vector<size_t> idx(blk.size());
    std::iota(idx.begin(), idx.end(), size_t(0));
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return blk[x].size() > blk[y].size(); });
    for (size_t i : idx) {
      size_t ln = 0;
      while (ln < lanes.size() && lanes[ln].size() + blk[i].size() > size_t(femdev::PL_TRIPS)) ++ln;
      if (ln == lanes.size()) lanes.emplace_back();
      lanes[ln].insert(lanes[ln].end(), blk[i].begin(), blk[i].end());
    }
    return true;
  };
  std::vector<int32_t> elem_local(E, -1);
  std::vector<uint8_t> in_wave(E, 0);
  std::vector<uint32_t> cur, cur_elems;
  std::vector<std::vector<std::vector<Item>>> cur_lanes;
  size_t cur_nlanes = 0;
  bool stage_fits = true;  // always: at most 16 x (13 x 9 + 1) values and 16 parity pads
  auto flush = [&]() {
    if (cur.empty()) return;
    std::sort(cur_elems.begin(), cur_elems.end());  // insertion order
    for (size_t i = 0; i < cur_elems.size(); ++i) elem_local[cur_elems[i]] = int32_t(i);
    size_t lane_base = pm.lanes.size();
    pm.lanes.resize(lane_base + 128, 0u);
    pm.lane_set.push_back(uint32_t(pm.n_waves));
    pm.nodes.resize(pm.nodes.size() + 2 * femdev::PL_NODES, 0u);
    uint32_t* nd = pm.nodes.data() + pm.nodes.size() - 2 * femdev::PL_NODES;
    size_t ln = 0;
    uint32_t trips = 0;
    // staging offset of each owned node's values: nodes whose CSR ranges
    // follow each other (a run) follow each other in the staging too, and a
    // run starts at the parity of its CSR position, so that 16-byte staging
    // chunks are 16-byte CSR chunks
    uint32_t lds = 0, prev_end = 0;
    for (size_t s = 0; s < cur.size(); ++s) {
      const uint32_t a = cur[s];
      const uint32_t base = uint32_t(row_ptr[6 * size_t(a)]);
      const uint32_t nb = uint32_t(bptr[a + 1] - bptr[a]);
      if (s == 0 || base != prev_end) lds += (lds ^ base) & 1u;
      nd[2 * s] = base;
      nd[2 * s + 1] = nb | lds << 18;
      lds += 13 * nb + 1;
      prev_end = base + 13 * nb + 1;
      for (const auto& lane : cur_lanes[s]) {
        uint64_t bits = 0;
        for (size_t t = 0; t < lane.size(); ++t) {
          const Item& it = lane[t];
          const uint64_t code = uint64_t(elem_local[it.e]) | uint64_t(it.p) << 5 | uint64_t(it.q) << 7 |
                                uint64_t(it.k) << 9;
          bits |= code << (13 * t);
        }
        bits |= uint64_t(lane.size()) << 52 | uint64_t(s) << 55 | uint64_t(nb) << 59;
        pm.lanes[lane_base + 2 * ln] = uint32_t(bits);
        pm.lanes[lane_base + 2 * ln + 1] = uint32_t(bits >> 32);
        trips = std::max(trips, uint32_t(lane.size()));
        ++ln;
      }
    }
    for (size_t i = 0; i < size_t(femdev::PL_ELEMS); ++i) {
      const bool used = i < cur_elems.size();
      const uint32_t e = used ? cur_elems[i] : 0u;
      pm.seid.push_back(used ? e : 0xFFFFFFFFu);
      for (int q = 0; q < 4; ++q) pm.sconn.push_back(used ? S.conn[4 * size_t(e) + q] : 0u);
      if (used) {
        elem_local[e] = -1;
        in_wave[e] = 0;
      }
    }
    stage_fits = stage_fits && lds <= uint32_t(femdev::PL_STAGE_VALUES);
    const uint32_t hdr = uint32_t(cur_elems.size()) | uint32_t(cur.size()) << 6 | trips << 11;
    for (int s = 0; s < femdev::PL_NODES; ++s) nd[2 * s + 1] |= hdr << 4;
    pm.n_waves++;
    cur.clear();
    cur_elems.clear();
    cur_lanes.clear();
    cur_nlanes = 0;
  };
  std::vector<std::vector<Item>> lanes;
  // greedy patches along a node order
  auto pack = [&](const std::vector<std::pair<uint64_t, uint32_t>>& ord) -> bool {
    pm.lanes.clear();
    pm.lane_set.clear();
    pm.nodes.clear();
    pm.seid.clear();
    pm.sconn.clear();
    pm.n_waves = 0;
    for (auto& o : ord) {
      const uint32_t a = o.second;
      if (!node_lanes(a, lanes)) return false;
      uint32_t new_elems = 0;
      for (uint32_t r = 0; r < deg[a]; ++r) new_elems += !in_wave[nel[nptr[a] + r] & 0x3FFFFFFFu];
      if (cur.size() == size_t(femdev::PL_NODES) || cur_nlanes + lanes.size() > 64 ||
          cur_elems.size() + new_elems > size_t(femdev::PL_ELEMS))
        flush();
      cur.push_back(a);
      cur_nlanes += lanes.size();
      cur_lanes.push_back(lanes);
      for (uint32_t r = 0; r < deg[a]; ++r) {
        const uint32_t e = nel[nptr[a] + r] & 0x3FFFFFFFu;
        if (!in_wave[e]) { in_wave[e] = 1; cur_elems.push_back(e); }
      }
    }
    flush();
    // one copy of each distinct descriptor set
    std::unordered_map<std::string, uint32_t> sets;
    std::vector<uint32_t> uniq;
    for (size_t w = 0; w < size_t(pm.n_waves); ++w) {
      const uint32_t* src = pm.lanes.data() + 128 * w;
      std::string key(reinterpret_cast<const char*>(src), 128 * sizeof(uint32_t));
      auto it = sets.find(key);
      if (it == sets.end()) {
        it = sets.emplace(std::move(key), uint32_t(uniq.size() / 128)).first;
        uniq.insert(uniq.end(), src, src + 128);
      }
      pm.lane_set[w] = it->second;
    }
    pm.lanes.swap(uniq);
    return true;
  };
  // Node-index order first: where consecutive nodes are neighbours (a
  // structured numbering) its patches are strips whose CSR values form one
  // contiguous range, written out as one run.  When that order scatters the
  // patches' elements (more than 5/4 of the fewest possible patches), the
  // spatial (Morton) order instead.
  std::vector<std::pair<uint64_t, uint32_t>> by_index(order);
  std::stable_sort(by_index.begin(), by_index.end(), [&](const std::pair<uint64_t, uint32_t>& x,
                                                         const std::pair<uint64_t, uint32_t>& y) {
    if (m->node_part.size() == m->nodes_number) {
      const uint32_t px = m->node_part[x.second], py = m->node_part[y.second];
      if (px != py) return px < py;
    }
    return x.second < y.second;
  });
  const size_t fewest = (order.size() + femdev::PL_NODES - 1) / femdev::PL_NODES;
  if (m->pl_order == 2) return pack(order) && stage_fits;
  if (!pack(by_index)) return false;
  if (m->pl_order == 0 && 4 * size_t(pm.n_waves) > 5 * fewest && !pack(order)) return false;
  return stage_fits;
}

// Corner-lane maps (plate_corners.hip).  Returns false when the mesh does not
// fit the kernel's exchange scheme (the caller then uses the block-lane maps).
bool build_plate_corner_maps(fem_model* m, const std::vector<uint64_t>& bptr, const std::vector<PairRef>& blocks,
                             const std::vector<int64_t>& row_ptr, femdev::PlateCornerMaps& pm) {
  const auto& S = m->set[PLATE];
  const size_t E = S.size();
  std::vector<uint32_t> deg, nptr, nel;
  std::vector<std::pair<uint64_t, uint32_t>> order;
  plate_node_incidence(m, deg, nptr, nel, order);
  for (const auto& o : order)
    if (deg[o.second] > 4 || bptr[o.second + 1] - bptr[o.second] > 31) return false;
  std::vector<uint32_t> cur, cur_elems;
  std::vector<int32_t> elem_local(E, -1);
  std::vector<uint8_t> in_patch(E, 0);
  auto block_of = [&](uint32_t a, uint32_t b) -> int {
    const PairRef* k0 = blocks.data() + bptr[a];
    const PairRef* k1 = blocks.data() + bptr[a + 1];
    const PairRef* kb = std::lower_bound(k0, k1, b, [](const PairRef& x, uint32_t v) { return x.b < v; });
    return (kb != k1 && kb->b == b) ? int(kb - k0) : -1;
  };
  // lane words of node a (4 lanes) into w[16]
  auto node_lanes = [&](uint32_t a, uint32_t* w) -> bool {
    const uint32_t d = deg[a];
    const uint32_t* el = &nel[nptr[a]];
    const int nb = int(bptr[a + 1] - bptr[a]);
    const int kself = block_of(a, a);
    if (kself < 0) return false;
    int kq[4][4], pj[4];
    std::vector<std::array<int, 2>> contrib(size_t(nb) * 2, {-1, -1});  // up to two (lane, q) per block
    std::vector<int> ccount(nb, 0);
    for (uint32_t j = 0; j < d; ++j) {
      const uint32_t e = el[j] & 0x3FFFFFFFu;
      pj[j] = int(el[j] >> 30);
      for (int q = 0; q < 4; ++q) {
        const int k = block_of(a, S.conn[4 * size_t(e) + q]);
        if (k < 0) return false;
        kq[j][q] = k;
        if (q == pj[j]) continue;
        if (ccount[k] == 2) return false;  // more than two elements share the pair
        contrib[2 * size_t(k) + ccount[k]++] = {int(j), q};
      }
    }
    for (int k = 0; k < nb; ++k)
      if (k != kself && !ccount[k]) return false;
    // shared blocks as edges between lanes; orient every path and cycle so that
    // each lane exports at most one block and completes at most one
    struct Edge { int u, v, k; };
    std::vector<Edge> edges;
    int ldeg[4] = {0, 0, 0, 0};
    for (int k = 0; k < nb; ++k)
      if (k != kself && ccount[k] == 2) {
        edges.push_back({contrib[2 * k][0], contrib[2 * k + 1][0], k});
        ldeg[contrib[2 * k][0]]++;
        ldeg[contrib[2 * k + 1][0]]++;
      }
    for (int j = 0; j < 4; ++j)
      if (ldeg[j] > 2) return false;
    int own_q[4] = {-1, -1, -1, -1}, prov_q[4] = {-1, -1, -1, -1}, src[4] = {0, 0, 0, 0};
    std::vector<char> used(edges.size(), 0);
    auto q_of = [&](int k, int j) { return contrib[2 * k][0] == j ? contrib[2 * k][1] : contrib[2 * k + 1][1]; };
    auto walk = [&](int v) -> bool {
      for (;;) {
        int i = -1;
        for (size_t x = 0; x < edges.size() && i < 0; ++x)
          if (!used[x] && (edges[x].u == v || edges[x].v == v)) i = int(x);
        if (i < 0) return true;
        used[i] = 1;
        const int w2 = edges[i].u == v ? edges[i].v : edges[i].u;
        if (own_q[v] >= 0 || prov_q[w2] >= 0) return false;
        own_q[v] = q_of(edges[i].k, v);
        src[v] = w2;
        prov_q[w2] = q_of(edges[i].k, w2);
        v = w2;
      }
    };
    for (int j = 0; j < 4; ++j)
      if (ldeg[j] == 1 && !walk(j)) return false;
    for (int j = 0; j < 4; ++j)
      if (!walk(j)) return false;
    for (uint32_t j = 0; j < 4; ++j) {
      uint32_t* o = w + 4 * j;
      if (j >= d) {
        o[0] = 0; o[1] = 127u; o[2] = 0; o[3] = 0;
        continue;
      }
      const int p = pj[j];
      int keep[3], nk = 0;
      for (int q = 0; q < 4; ++q)
        if (q != p && q != own_q[j] && q != prov_q[j]) keep[nk++] = q;
      int qs[4] = {p, 0, 0, 0}, u = 0;
      qs[1] = prov_q[j] >= 0 ? prov_q[j] : keep[u++];
      qs[2] = own_q[j] >= 0 ? own_q[j] : keep[u++];
      qs[3] = keep[u++];
      if (u != nk) return false;
      const uint32_t e = el[j] & 0x3FFFFFFFu;
      o[0] = uint32_t(row_ptr[6 * size_t(a)]);
      o[1] = uint32_t(elem_local[e]) | uint32_t(p) << 7 | uint32_t(qs[1]) << 9 | uint32_t(qs[2]) << 11 |
             uint32_t(qs[3]) << 13 | uint32_t(nb) << 15 | (j == 0 ? d : 0u) << 20 | uint32_t(j == 0) << 23 |
             uint32_t(prov_q[j] < 0) << 24 | uint32_t(own_q[j] >= 0) << 25 | uint32_t(src[j]) << 26;
      o[2] = uint32_t(kq[j][qs[0]]) | uint32_t(kq[j][qs[1]]) << 5 | uint32_t(kq[j][qs[2]]) << 10 |
             uint32_t(kq[j][qs[3]]) << 15;
      o[3] = 0;
    }
    return true;
  };
  auto flush = [&]() -> bool {
    if (cur.empty()) return true;
    std::sort(cur.begin(), cur.end());
    std::sort(cur_elems.begin(), cur_elems.end());  // insertion order
    for (size_t i = 0; i < cur_elems.size(); ++i) elem_local[cur_elems[i]] = int32_t(i);
    const uint32_t elem_off = uint32_t(pm.peid.size());
    const size_t lane0 = pm.lane.size();
    pm.lane.resize(lane0 + 4 * size_t(4 * femdev::PC_NODES), 0);
    for (size_t l = 0; l < size_t(4 * femdev::PC_NODES); ++l) pm.lane[lane0 + 4 * l + 1] = 127u;
    for (size_t n = 0; n < cur.size(); ++n)
      if (!node_lanes(cur[n], &pm.lane[lane0 + 16 * n])) return false;
    for (uint32_t e : cur_elems) {
      pm.peid.push_back(e);
      for (int q = 0; q < 4; ++q) pm.pconn.push_back(S.conn[4 * size_t(e) + q]);
      elem_local[e] = -1;
    }
    pm.hdr.insert(pm.hdr.end(), {elem_off, uint32_t(cur_elems.size())});
    pm.n_patches++;
    cur.clear();
    cur_elems.clear();
    return true;
  };
  for (auto& o : order) {
    const uint32_t a = o.second;
    uint32_t new_elems = 0;
    for (uint32_t r = 0; r < deg[a]; ++r) new_elems += !in_patch[nel[nptr[a] + r] & 0x3FFFFFFFu];
    if (cur.size() == size_t(femdev::PC_NODES) || cur_elems.size() + new_elems > size_t(femdev::PC_MAX_ELEMS)) {
      for (uint32_t e : cur_elems) in_patch[e] = 0;
      if (!flush()) return false;
    }
    cur.push_back(a);
    for (uint32_t r = 0; r < deg[a]; ++r) {
      const uint32_t e = nel[nptr[a] + r] & 0x3FFFFFFFu;
      if (!in_patch[e]) { in_patch[e] = 1; cur_elems.push_back(e); }
    }
  }
  for (uint32_t e : cur_elems) in_patch[e] = 0;
  return flush();
}

std::string build_pattern(fem_model* m) {
  std::string e;
  const size_t N = m->nodes_number;
  const int64_t n_rows = int64_t(N) * 6;
  // element classes (type, rzero) -> per pair masks
  std::unordered_map<uint32_t, uint16_t> class_id;
  femdev::GenericMaps g;
  auto get_class = [&](int t, uint16_t rz) -> uint16_t {
    uint32_t key = uint32_t(t) << 16 | rz;
    auto it = class_id.find(key);
    if (it != class_id.end()) return it->second;
    uint16_t id = uint16_t(class_id.size());
    class_id[key] = id;
    int nn = femdev::NODES_PER[t];
    uint32_t off = 0;
    for (int pq = 0; pq < 16; ++pq) {
      Mask mk = 0;
      if (pq < nn * nn) mk = element_pair_mask(t, rz, pq / nn, pq % nn);
      g.class_mask.push_back(mk);
      g.class_off.push_back(off);
      off += uint32_t(__builtin_popcountll(mk));
    }
    return id;
  };
  // element visiting order = global insertion order
  struct ERef {
    uint64_t seq;
    int t;
    uint32_t pos;
  };
  std::vector<ERef> order;
  for (int t = 0; t < 3; ++t)
    for (uint32_t i = 0; i < m->set[t].size(); ++i) order.push_back({m->set[t].seq[i], t, i});
  std::sort(order.begin(), order.end(), [](const ERef& a, const ERef& b) { return a.seq < b.seq; });
  for (int t = 0; t < 3; ++t) {
    g.elem_class[t].resize(m->set[t].size());
    g.elem_base[t].resize(m->set[t].size());
  }
  int64_t ncontrib = 0;
  std::vector<uint32_t> deg(N + 1, 0);
  for (auto& r : order) {
    uint16_t rz = m->rzero[r.t][r.pos];
    if (rz & 0x8000) return "Zero length vector could not be aligned!";
    uint16_t cls = get_class(r.t, rz);
    g.elem_class[r.t][r.pos] = cls;
    g.elem_base[r.t][r.pos] = ncontrib;
    ncontrib += g.class_off[cls * 16 + 15] + __builtin_popcountll(g.class_mask[cls * 16 + 15]);
    int nn = femdev::NODES_PER[r.t];
    for (int p = 0; p < nn; ++p) deg[m->set[r.t].conn[size_t(r.pos) * nn + p]] += nn;
  }
  g.n_contrib = ncontrib;
  if (ncontrib > int64_t(UINT32_MAX)) return "too many element contributions for 32-bit slot maps";
  // node -> (neighbour, mask) lists, merged
  std::vector<uint64_t> nptr(N + 1, 0);
  for (size_t a = 0; a < N; ++a) nptr[a + 1] = nptr[a] + deg[a];
  std::vector<PairRef> pr(nptr[N]);
  std::vector<uint64_t> fill(nptr.begin(), nptr.end() - 1);
  for (auto& r : order) {
    int nn = femdev::NODES_PER[r.t];
    const uint32_t* c = &m->set[r.t].conn[size_t(r.pos) * nn];
    uint16_t cls = g.elem_class[r.t][r.pos];
    for (int p = 0; p < nn; ++p)
      for (int q = 0; q < nn; ++q) pr[fill[c[p]]++] = {c[q], g.class_mask[cls * 16 + p * nn + q]};
  }
  std::vector<uint64_t> bptr(N + 1, 0);
  std::vector<PairRef> blocks;
  blocks.reserve(pr.size() / 2 + 1);
  bool single_pairs = true;  // every off-diagonal node pair belongs to one element
  for (size_t a = 0; a < N; ++a) {
    auto* b0 = pr.data() + nptr[a];
    auto* b1 = pr.data() + nptr[a + 1];
    std::stable_sort(b0, b1, [](const PairRef& x, const PairRef& y) { return x.b < y.b; });
    for (auto* it = b0; it != b1;) {
      PairRef acc = *it;
      int cnt = 1;
      for (++it; it != b1 && it->b == acc.b; ++it) {
        acc.mask |= it->mask;
        cnt++;
      }
      if (cnt > 1 && acc.b != uint32_t(a)) single_pairs = false;
      if (acc.mask) blocks.push_back(acc);
    }
    bptr[a + 1] = blocks.size();
  }
  pr.clear();
  pr.shrink_to_fit();
  // rows: length and per-block row offsets
  std::vector<int64_t> row_ptr(static_cast<size_t>(n_rows) + 1, 0);
  std::vector<uint16_t> boff(blocks.size() * 6);
  for (size_t a = 0; a < N; ++a)
    for (int i = 0; i < 6; ++i) {
      int64_t len = 0;
      for (uint64_t k = bptr[a]; k < bptr[a + 1]; ++k) {
        if (len > 65535) return "row too long for the slot map";
        boff[k * 6 + i] = uint16_t(len);
        len += __builtin_popcountll((blocks[k].mask >> (6 * i)) & 63ull);
      }
      row_ptr[6 * a + i + 1] = len;
    }
  for (int64_t r = 0; r < n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
  const int64_t nnz = row_ptr[n_rows];
  if (nnz > int64_t(INT32_MAX) * 4) return "pattern too large";
  std::vector<int32_t> col(static_cast<size_t>(nnz));
  for (size_t a = 0; a < N; ++a)
    for (int i = 0; i < 6; ++i) {
      int64_t pos = row_ptr[6 * a + i];
      for (uint64_t k = bptr[a]; k < bptr[a + 1]; ++k)
        for (int j = 0; j < 6; ++j)
          if (bit(blocks[k].mask, i, j)) col[size_t(pos++)] = int32_t(6 * int64_t(blocks[k].b) + j);
    }
  e = femdev::upload_pattern(m->dev, n_rows, row_ptr.data(), nnz, col.data());
  if (!e.empty()) return e;
  m->nnz = nnz;
  m->row_ptr_host = row_ptr;
  m->adj_valid = false;
  m->pat_bptr = bptr;
  m->pat_bnode.resize(blocks.size());
  m->pat_bmask.resize(blocks.size());
  for (size_t k = 0; k < blocks.size(); ++k) {
    m->pat_bnode[k] = blocks[k].b;
    m->pat_bmask[k] = blocks[k].mask;
  }
  m->pat_boff = boff;
  col.clear();
  col.shrink_to_fit();
  {
    femdev::PlateMaps pm;
    if (!m->force_generic && build_plate_maps(m, bptr, blocks, row_ptr, pm)) {
      femdev::PlateCornerMaps pcm;
      femdev::PlateBlockMaps pbm;
      femdev::PlateLaneMaps plm;
      if (m->pf_kernel == 3 && build_plate_lane_maps(m, bptr, blocks, row_ptr, plm)) {
        double mat[3];
        e = femdev::set_plate_lane_material(m->dev, plate_material_uniform(m, mat), mat);
        if (!e.empty()) return e;
        e = femdev::upload_plate_lane_maps(m->dev, plm);
        if (!e.empty()) return e;
        m->path = "plate_lanes";
      } else if (m->pf_kernel == 2 && build_plate_corner_maps(m, bptr, blocks, row_ptr, pcm)) {
        e = femdev::upload_plate_corner_maps(m->dev, pcm, m->pf_wg_per_cu < 0 ? 0 : m->pf_wg_per_cu);
        if (!e.empty()) return e;
        m->path = "plate_corners";
      } else if (m->pf_kernel >= 1 && build_plate_block_maps(m, bptr, blocks, row_ptr, pbm)) {
        e = femdev::upload_plate_block_maps(m->dev, pbm, m->pf_wg_per_cu < 0 ? 0 : m->pf_wg_per_cu);
        if (!e.empty()) return e;
        m->path = "plate_blocks";
      } else {
        e = femdev::upload_plate_maps(m->dev, pm, m->pf_wg_per_cu < 0 ? 2 : m->pf_wg_per_cu);
        if (!e.empty()) return e;
        m->path = "plate_fused";
      }
      m->pattern_valid = true;
      return "";
    }
  }
  if (!m->force_generic && single_pairs && !m->set[PLATE].size() && m->set[TRUSS].size() + m->set[BEAM].size() &&
      m->set[TRUSS].size() < (1u << 30) && m->set[BEAM].size() < (1u << 30)) {
    femdev::BarMaps bm;
    bool ok = true;
    for (int t = 0; t < 2 && ok; ++t) {
      const auto& S = m->set[t];
      bm.bid[t].resize(S.size() * 4);
      for (size_t e = 0; e < S.size() && ok; ++e)
        for (int pq = 0; pq < 4; ++pq) {
          const uint32_t a = S.conn[2 * e + pq / 2], b = S.conn[2 * e + pq % 2];
          auto* k0 = blocks.data() + bptr[a];
          auto* k1 = blocks.data() + bptr[a + 1];
          auto* kb = std::lower_bound(k0, k1, b, [](const PairRef& x, uint32_t v) { return x.b < v; });
          if (kb == k1 || kb->b != b) { ok = false; break; }
          bm.bid[t][4 * e + pq] = uint32_t(kb - blocks.data());
        }
    }
    if (ok && blocks.size() < (size_t(1) << 32)) {
      bm.bmask.resize(blocks.size());
      for (size_t k = 0; k < blocks.size(); ++k) bm.bmask[k] = blocks[k].mask;
      bm.boff = boff;
      bm.inc_ptr.assign(N + 1, 0);
      for (auto& r : order) {
        const uint32_t* c = &m->set[r.t].conn[size_t(r.pos) * 2];
        bm.inc_ptr[c[0] + 1]++;
        bm.inc_ptr[c[1] + 1]++;
      }
      for (size_t a = 0; a < N; ++a) bm.inc_ptr[a + 1] += bm.inc_ptr[a];
      bm.inc.resize(bm.inc_ptr[N]);
      std::vector<uint32_t> f(bm.inc_ptr.begin(), bm.inc_ptr.end() - 1);
      for (auto& r : order) {
        const uint32_t* c = &m->set[r.t].conn[size_t(r.pos) * 2];
        for (int p = 0; p < 2; ++p) bm.inc[f[c[p]]++] = uint32_t(r.t) << 31 | r.pos << 1 | uint32_t(p);
      }
      // workgroup ranges of consecutive nodes.  Nodes whose incident elements
      // are all trusses go to the truss kernel (<= 85 nodes, <= 512 entries);
      // the others to the bar kernel (<= 40 nodes, <= 256 entries, <= 5120
      // CSR values, since it stages the rows in LDS).
      auto truss_only = [&](uint32_t a) {
        for (uint32_t k = bm.inc_ptr[a]; k < bm.inc_ptr[a + 1]; ++k)
          if (bm.inc[k] >> 31) return false;
        return true;
      };
      std::vector<uint32_t> truss_ranges;
      uint32_t a0 = 0;
      while (a0 < N && ok) {
        const bool tonly = truss_only(a0);
        const uint32_t max_nodes = uint32_t(tonly ? femdev::TRUSS_WG_NODES : femdev::BAR_WG_NODES);
        const uint32_t max_entries = uint32_t(tonly ? femdev::TRUSS_WG_ENTRIES : femdev::BAR_WG_ENTRIES);
        uint32_t na = 0;
        while (a0 + na < N && na < max_nodes && (na == 0 || truss_only(a0 + na) == tonly) &&
               bm.inc_ptr[a0 + na + 1] - bm.inc_ptr[a0] <= max_entries &&
               (tonly || row_ptr[6 * size_t(a0 + na + 1)] - row_ptr[6 * size_t(a0)] <= int64_t(femdev::BAR_WG_VALUES)))
          na++;
        if (na == 0) ok = false;  // one node with too many incident elements or CSR values
        auto& dst = tonly ? truss_ranges : bm.ranges;
        dst.push_back(a0);
        dst.push_back(na);
        a0 += na;
      }
      bm.n_bar_ranges = bm.ranges.size() / 2;
      bm.slot_ptr.assign(1, 0);
      for (size_t r = 0; r < bm.n_bar_ranges; ++r) {
        const uint32_t r0 = bm.ranges[2 * r], rn = bm.ranges[2 * r + 1];
        for (uint32_t k = bm.inc_ptr[r0]; k < bm.inc_ptr[r0 + rn]; ++k) bm.slot_inc.push_back(k);
        bm.slot_ptr.push_back(uint32_t(bm.slot_inc.size()));
      }
      bm.ranges.insert(bm.ranges.end(), truss_ranges.begin(), truss_ranges.end());
    }
    if (ok && blocks.size() < (size_t(1) << 32)) {
      e = femdev::upload_bar_maps(m->dev, bm);
      if (!e.empty()) return e;
      m->path = "bar_owner";
      m->pattern_valid = true;
      return "";
    }
  }
  // slot of every contribution, then gather lists in insertion order
  std::vector<uint32_t> slot_of(static_cast<size_t>(ncontrib));
  for (auto& r : order) {
    int nn = femdev::NODES_PER[r.t];
    const uint32_t* c = &m->set[r.t].conn[size_t(r.pos) * nn];
    uint16_t cls = g.elem_class[r.t][r.pos];
    int64_t base = g.elem_base[r.t][r.pos];
    for (int p = 0; p < nn; ++p)
      for (int q = 0; q < nn; ++q) {
        int pq = p * nn + q;
        Mask mk = g.class_mask[cls * 16 + pq];
        if (!mk) continue;
        uint32_t a = c[p], b = c[q];
        auto* k0 = blocks.data() + bptr[a];
        auto* k1 = blocks.data() + bptr[a + 1];
        auto* kb = std::lower_bound(k0, k1, b, [](const PairRef& x, uint32_t v) { return x.b < v; });
        size_t k = size_t(kb - blocks.data());
        Mask um = blocks[k].mask;
        int64_t ci = base + g.class_off[cls * 16 + pq];
        for (int i = 0; i < 6; ++i)
          for (int j = 0; j < 6; ++j)
            if (bit(mk, i, j)) {
              int below = __builtin_popcountll(((um >> (6 * i)) & 63ull) & ((1ull << j) - 1));
              slot_of[size_t(ci++)] = uint32_t(row_ptr[6 * a + i] + boff[k * 6 + i] + below);
            }
      }
  }
  g.slot_ptr.assign(size_t(nnz) + 1, 0);
  for (uint32_t s : slot_of) g.slot_ptr[s + 1] += 1;
  for (int64_t v = 0; v < nnz; ++v) g.slot_ptr[v + 1] += g.slot_ptr[v];
  g.slot_src.resize(size_t(ncontrib));
  {
    std::vector<int64_t> f(g.slot_ptr.begin(), g.slot_ptr.end() - 1);
    for (int64_t c = 0; c < ncontrib; ++c) g.slot_src[size_t(f[slot_of[size_t(c)]]++)] = uint32_t(c);
  }
  e = femdev::upload_generic(m->dev, g);
  if (!e.empty()) return e;
  m->path = "generic";
  m->pattern_valid = true;
  return "";
}

// upload + device setup + pattern, as needed
int ensure_ready(fem_model* m) {
  if (!m->dev) {
    std::string err;
    m->dev = femdev::create_device(err);
    if (!m->dev) return fail(m, FEM_ERR_DEVICE, err);
  }
  if (m->dirty) {
    auto e = femdev::upload_nodes(m->dev, m->xyz.data(), m->node_number.size());
    for (int t = 0; t < 3 && e.empty(); ++t) {
      e = femdev::upload_elements(m->dev, t, m->set[t].size(), m->set[t].conn.data(), m->set[t].params.data());
      if (e.empty()) e = femdev::run_setup(m->dev, t, m->set[t].size(), m->rel_tol, m->abs_tol, m->rzero[t]);
    }
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
    m->dirty = false;
    m->pattern_valid = false;
    m->own_vals = false;
  }
  if (!m->pattern_valid) {
    // a new pattern comes with a new (unwritten) values buffer
    m->own_vals = false;
    auto e = build_pattern(m);
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  }
  return FEM_OK;
}

// adjoint maps from the pattern's node-pair blocks, built on first use
int ensure_adjoint(fem_model* m) {
  int rc = ensure_ready(m);
  if (rc) return rc;
  if (m->adj_valid) return FEM_OK;
  femdev::AdjointMaps am;
  const size_t N = m->nodes_number;
  for (int t = 0; t < 3; ++t) {
    const int nn = femdev::NODES_PER[t];
    const auto& S = m->set[t];
    am.bid[t].resize(S.size() * nn * nn);
    for (size_t e = 0; e < S.size(); ++e)
      for (int p = 0; p < nn; ++p)
        for (int q = 0; q < nn; ++q) {
          const uint32_t a = S.conn[e * nn + p], b = S.conn[e * nn + q];
          const uint32_t* k0 = m->pat_bnode.data() + m->pat_bptr[a];
          const uint32_t* k1 = m->pat_bnode.data() + m->pat_bptr[a + 1];
          const uint32_t* kb = std::lower_bound(k0, k1, b);
          if (kb == k1 || *kb != b) return fail(m, FEM_ERR_STATE, "element node pair missing from the pattern");
          am.bid[t][(e * nn + p) * nn + q] = uint32_t(kb - m->pat_bnode.data());
        }
  }
  if (m->pat_bnode.size() >= (size_t(1) << 32)) return fail(m, FEM_ERR_ARGUMENT, "too many node-pair blocks");
  am.flat = !m->set[TRUSS].size() && !m->set[BEAM].size() && m->set[PLATE].size();
  for (uint16_t rz : m->rzero[PLATE]) am.flat = am.flat && (rz & 0x4000);
  if (am.flat) {
    const Mask MOFF = plate_off_mask(), MSELF = plate_off_mask() | (1ull << 35);
    const auto& S = m->set[PLATE];
    for (size_t e = 0; e < S.size() && am.flat; ++e)
      for (int pq = 0; pq < 16; ++pq)
        if (m->pat_bmask[am.bid[PLATE][16 * e + pq]] != ((pq / 4 == pq % 4) ? MSELF : MOFF)) am.flat = false;
    // the fixed in-row layout needs 13 nb + 1 values per node, nb < 64 and
    // 32-bit CSR positions
    if (m->row_ptr_host.back() >= (int64_t(1) << 32)) am.flat = false;
    for (size_t a = 0; a < N && am.flat; ++a) {
      const uint64_t nb = m->pat_bptr[a + 1] - m->pat_bptr[a];
      if (nb && (nb > 63 || m->row_ptr_host[6 * a + 6] - m->row_ptr_host[6 * a] != int64_t(13 * nb + 1)))
        am.flat = false;
    }
    if (am.flat) {
      am.kpack.resize(S.size() * 8);
      for (size_t e = 0; e < S.size(); ++e)
        for (int p = 0; p < 4; ++p) {
          const uint32_t a = S.conn[4 * e + p];
          const uint64_t k0 = m->pat_bptr[a];
          uint32_t w = uint32_t(m->pat_bptr[a + 1] - k0) << 24;
          for (int q = 0; q < 4; ++q) w |= uint32_t(am.bid[PLATE][16 * e + 4 * p + q] - k0) << (6 * q);
          am.kpack[8 * e + 2 * p] = uint32_t(m->row_ptr_host[6 * size_t(a)]);
          am.kpack[8 * e + 2 * p + 1] = w;
        }
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (size_t a = 0; a < N; ++a)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], m->xyz[3 * a + k]);
          hi[k] = std::max(hi[k], m->xyz[3 * a + k]);
        }
      // Morton order of the element centroids on a 2^21 grid over the bounding
      // box (a regular 2^k x 2^k mesh then falls into aligned 4 x 4 patches)
      std::vector<std::pair<uint64_t, uint32_t>> ord(S.size());
      for (size_t e = 0; e < S.size(); ++e) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
          double c = 0.0;
          for (int p = 0; p < 4; ++p) c += m->xyz[3 * size_t(S.conn[4 * e + p]) + k];
          const double span = hi[k] - lo[k];
          const double qd = span > 0 ? (0.25 * c - lo[k]) / span * 2097152.0 : 0.0;
          code |= spread3(uint64_t(std::min(std::max(qd, 0.0), 2097151.0))) << k;
        }
        ord[e] = {code, uint32_t(e)};
      }
      std::sort(ord.begin(), ord.end());
      am.eorder.resize(S.size());
      for (size_t i = 0; i < S.size(); ++i) am.eorder[i] = ord[i].second;
    }
  }
  am.bmask = m->pat_bmask;
  am.boff = m->pat_boff;
  // per node: (element, node) gradient rows of trusses and plates in insertion order
  am.xoff[TRUSS] = 0;
  am.xoff[BEAM] = int64_t(m->set[TRUSS].size()) * 2;
  am.xoff[PLATE] = am.xoff[BEAM];
  am.n_xs_rows = am.xoff[PLATE] + int64_t(m->set[PLATE].size()) * 4;
  if (am.n_xs_rows >= (int64_t(1) << 32)) return fail(m, FEM_ERR_ARGUMENT, "too many element nodes");
  struct ERef {
    uint64_t seq;
    int t;
    uint32_t pos;
  };
  std::vector<ERef> order;
  for (int t : {TRUSS, PLATE})
    for (uint32_t i = 0; i < m->set[t].size(); ++i) order.push_back({m->set[t].seq[i], t, i});
  std::sort(order.begin(), order.end(), [](const ERef& x, const ERef& y) { return x.seq < y.seq; });
  am.inc_ptr.assign(N + 1, 0);
  for (auto& r : order) {
    const int nn = femdev::NODES_PER[r.t];
    for (int p = 0; p < nn; ++p) am.inc_ptr[m->set[r.t].conn[size_t(r.pos) * nn + p] + 1]++;
  }
  for (size_t a = 0; a < N; ++a) am.inc_ptr[a + 1] += am.inc_ptr[a];
  am.xpos.assign(size_t(am.n_xs_rows), 0);
  {
    std::vector<uint32_t> f(am.inc_ptr.begin(), am.inc_ptr.end() - 1);
    for (auto& r : order) {
      const int nn = femdev::NODES_PER[r.t];
      for (int p = 0; p < nn; ++p)
        am.xpos[size_t(am.xoff[r.t] + int64_t(r.pos) * nn + p)] = f[m->set[r.t].conn[size_t(r.pos) * nn + p]]++;
    }
  }
  auto e = femdev::upload_adjoint_maps(m->dev, am);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  m->adj_valid = true;
  return FEM_OK;
}

int run_backward(fem_model* m, int mode, const double* g, double* dp_truss, double* dp_beam, double* dp_plate,
                 double* dxyz, void* stream) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!g) return fail(m, FEM_ERR_ARGUMENT, mode == 0 ? "dL/dK values are required" : "displacements are required");
  if (dxyz && m->set[BEAM].size())
    return fail(m, FEM_ERR_ARGUMENT,
                "node-coordinate gradients are not available for models with beam elements: the section roll "
                "angle (beam.rs:185-202, acos of a sign-blind cosine) is not differentiable at its usual values 0 and pi");
  int rc = ensure_adjoint(m);
  if (rc) return rc;
  double* dp[3] = {dp_truss, dp_beam, dp_plate};
  auto e = femdev::run_adjoint(m->dev, mode, g, dp, dxyz, stream, m->wide_loads);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

// reruns the setup kernels after a parameter or coordinate change; a changed
// frame zero pattern invalidates the sparsity pattern
int rerun_setup(fem_model* m, bool all_types, int type) {
  for (int t = 0; t < 3; ++t) {
    if (!all_types && t != type) continue;
    std::vector<uint16_t> rz;
    auto e = femdev::run_setup(m->dev, t, m->set[t].size(), m->rel_tol, m->abs_tol, rz);
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
    if (rz != m->rzero[t]) {
      m->rzero[t] = rz;
      m->pattern_valid = false;
      m->adj_valid = false;
    }
  }
  return FEM_OK;
}

// ------------------------------------------------------------------ batches
// true when every plate frame of m is exactly the identity: the normal
// (X4 - X3) x (X2 - X3) of quad_rotation has zero x and y components and a
// positive z component, which the minimal rotation maps to I exactly
bool plates_flat_identity(const fem_model* m) {
  const auto& S = m->set[PLATE];
  for (size_t e = 0; e < S.size(); ++e) {
    const double* p2 = &m->xyz[3 * size_t(S.conn[4 * e + 1])];
    const double* p3 = &m->xyz[3 * size_t(S.conn[4 * e + 2])];
    const double* p4 = &m->xyz[3 * size_t(S.conn[4 * e + 3])];
    const double e34[3] = {p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
    const double e32[3] = {p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
    const double nx = e34[1] * e32[2] - e34[2] * e32[1], ny = e34[2] * e32[0] - e34[0] * e32[2];
    const double nz = e34[0] * e32[1] - e34[1] * e32[0];
    if (!(nx == 0.0 && ny == 0.0 && nz > 0.0)) return false;
  }
  return true;
}

// appends src's nodes and elements to dst (indices shifted; element order kept)
void append_model(fem_model* dst, const fem_model* src, uint32_t part) {
  const uint32_t off = uint32_t(dst->node_number.size());
  for (size_t i = 0; i < src->node_number.size(); ++i) {
    const uint32_t idx = off + uint32_t(i);
    dst->node_idx[idx + 1] = idx;
    dst->node_number.push_back(idx + 1);
    dst->node_plates.emplace_back();
    dst->node_part.push_back(part);
  }
  dst->xyz.insert(dst->xyz.end(), src->xyz.begin(), src->xyz.end());
  struct ERef {
    uint64_t seq;
    int t;
    uint32_t pos;
  };
  std::vector<ERef> order;
  for (int t = 0; t < 3; ++t)
    for (uint32_t i = 0; i < src->set[t].size(); ++i) order.push_back({src->set[t].seq[i], t, i});
  std::sort(order.begin(), order.end(), [](const ERef& x, const ERef& y) { return x.seq < y.seq; });
  for (auto& r : order) {
    const int nn = femdev::NODES_PER[r.t], np = femdev::PARAMS_PER[r.t];
    auto& D = dst->set[r.t];
    const auto& S = src->set[r.t];
    const uint32_t num = uint32_t(D.size()) + 1;
    D.by_number[num] = uint32_t(D.size());
    D.number.push_back(num);
    for (int p = 0; p < nn; ++p) D.conn.push_back(S.conn[size_t(r.pos) * nn + p] + off);
    D.params.insert(D.params.end(), S.params.begin() + size_t(r.pos) * np, S.params.begin() + size_t(r.pos + 1) * np);
    D.seq.push_back(dst->seq++);
  }
  dst->dirty = true;
  dst->pattern_valid = false;
}

int batch_fail(fem_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}

int batch_ready(fem_batch* b) {
  if (b->ready) return FEM_OK;
  for (auto* g : b->groups) {
    int rc = ensure_ready(g);
    if (rc) return batch_fail(b, rc, g->err);
  }
  for (auto& mm : b->meshes) {
    const fem_model* g = b->groups[size_t(mm.group)];
    const size_t r0 = 6 * mm.node_off, r1 = 6 * (mm.node_off + mm.n_nodes);
    mm.val_off = g->row_ptr_host[r0];
    mm.nnz = g->row_ptr_host[r1] - mm.val_off;
  }
  b->ready = true;
  return FEM_OK;
}

std::string launch_values(fem_model* g, double* val, void* stream, bool timed) {
  return g->path == "plate_fused" ? femdev::assemble_plate_fused(g->dev, val, stream, timed)
         : g->path == "plate_lanes" ? femdev::assemble_plate_lanes(g->dev, val, stream, timed)
         : g->path == "plate_blocks" ? femdev::assemble_plate_blocks(g->dev, val, stream, timed)
         : g->path == "plate_corners" ? femdev::assemble_plate_corners(g->dev, val, stream, timed)
         : g->path == "bar_owner" ? femdev::assemble_bar_owner(g->dev, val, stream, timed)
                                  : femdev::assemble_generic(g->dev, val, stream, timed);
}

std::string batch_launch_groups_one(fem_model* g, void* stream) {
  return launch_values(g, femdev::values(g->dev), stream, false);
}

// side streams and fork / join events for groups 1.. (created before a capture)
std::string batch_ensure_branches(fem_batch* b) {
  const size_t ng = b->groups.size();
  while (b->side.size() + 1 < ng) {
    void *s = nullptr, *f = nullptr, *j = nullptr;
    auto e = femdev::stream_create(s);
    if (e.empty()) e = femdev::event_create(f);
    if (e.empty()) e = femdev::event_create(j);
    if (!e.empty()) {
      femdev::stream_destroy(s);
      femdev::event_destroy(f);
      femdev::event_destroy(j);
      return e;
    }
    b->side.push_back(s);
    b->fork_ev.push_back(f);
    b->join_ev.push_back(j);
  }
  return "";
}

// the groups' launches as parallel branches of one captured graph, or one
// after another (schedule 1)
std::string batch_launch_branches(fem_batch* b) {
  const size_t ng = b->groups.size();
  if (b->schedule == 1) {
    for (size_t k = 0; k < ng; ++k) {
      auto* g = b->groups[k];
      auto e = launch_values(g, femdev::values(g->dev), b->stream, false);
      if (!e.empty()) return e;
    }
    return "";
  }
  for (size_t k = 1; k < ng; ++k) {
    auto e = femdev::stream_wait(b->side[k - 1], b->stream, b->fork_ev[k - 1]);
    if (!e.empty()) return e;
  }
  for (size_t k = 0; k < ng; ++k) {
    auto* g = b->groups[k];
    auto e = launch_values(g, femdev::values(g->dev), k ? b->side[k - 1] : b->stream, false);
    if (!e.empty()) return e;
  }
  for (size_t k = 1; k < ng; ++k) {
    auto e = femdev::stream_wait(b->stream, b->side[k - 1], b->join_ev[k - 1]);
    if (!e.empty()) return e;
  }
  return "";
}

int batch_capture(fem_batch* b) {
  femdev::graph_destroy(b->graph);
  b->graph = nullptr;
  auto e = batch_ensure_branches(b);
  if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
  e = femdev::graph_begin(b->stream);
  if (e.empty()) {
    e = batch_launch_branches(b);
    femdev::Graph* g = nullptr;
    auto e2 = femdev::graph_end(b->stream, g);
    if (e.empty()) e = e2;
    b->graph = e.empty() ? g : nullptr;
    if (!e.empty()) femdev::graph_destroy(g);
  }
  return e.empty() ? FEM_OK : batch_fail(b, FEM_ERR_DEVICE, e);
}

// median of `reps` timed launches of one group on the batch stream
int time_group(fem_batch* b, fem_model* g, int reps, double& ms) {
  std::vector<double> t;
  for (int r = -1; r < reps; ++r) {
    auto e = femdev::timer_start(b->timer, b->stream);
    if (e.empty()) e = batch_launch_groups_one(g, b->stream);
    double x = 0;
    if (e.empty()) e = femdev::timer_stop(b->timer, b->stream, x);
    if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
    if (r >= 0) t.push_back(x);
  }
  std::sort(t.begin(), t.end());
  ms = t[t.size() / 2];
  return FEM_OK;
}

// median of `reps` timed replays of the batch's graph (one warm-up)
int time_graph(fem_batch* b, int reps, double& ms) {
  std::vector<double> t;
  for (int r = -1; r < reps; ++r) {
    auto e = femdev::timer_start(b->timer, b->stream);
    if (e.empty()) e = femdev::graph_launch(b->graph, b->stream);
    double x = 0;
    if (e.empty()) e = femdev::timer_stop(b->timer, b->stream, x);
    if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
    if (r >= 0) t.push_back(x);
  }
  std::sort(t.begin(), t.end());
  ms = t[t.size() / 2];
  return FEM_OK;
}

// Persistent autotune cache: one line per (gfx arch, batch signature),
// "arch signature schedule pn,wg,bs;pn,wg,bs;...", in the file named by
// FEM_AMD_TUNE_CACHE or else $XDG_CACHE_HOME/fem_amd/tune.txt
// ($HOME/.cache/fem_amd/tune.txt).  Unwritable locations are skipped.
std::string tune_cache_path() {
  if (const char* p = std::getenv("FEM_AMD_TUNE_CACHE")) return p;
  std::string base;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) base = x;
  else if (const char* h = std::getenv("HOME")) base = std::string(h) + "/.cache";
  else return "";
  return base + "/fem_amd/tune.txt";
}

bool tune_cache_find(const std::string& key, int& schedule, std::vector<std::array<int, 3>>& cfg) {
  const std::string path = tune_cache_path();
  if (path.empty()) return false;
  std::ifstream in(path);
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {  // the last entry of a key wins
    std::istringstream ls(line);
    std::string arch, sig, cfgs;
    int sched = 0;
    if (!(ls >> arch >> sig >> sched >> cfgs) || arch + " " + sig != key) continue;
    std::vector<std::array<int, 3>> c;
    std::istringstream cs(cfgs);
    std::string g;
    bool ok = true;
    while (std::getline(cs, g, ';')) {
      std::array<int, 3> v{};
      if (std::sscanf(g.c_str(), "%d,%d,%d", &v[0], &v[1], &v[2]) != 3) { ok = false; break; }
      c.push_back(v);
    }
    if (ok) { schedule = sched; cfg = c; found = true; }
  }
  return found;
}

void tune_cache_store(const std::string& key, int schedule, const std::vector<std::array<int, 3>>& cfg) {
  const std::string path = tune_cache_path();
  if (path.empty()) return;
  for (size_t i = 1; i < path.size(); ++i)  // parent directories
    if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
  std::ofstream out(path, std::ios::app);
  if (!out) return;
  out << key << ' ' << schedule << ' ';
  for (size_t k = 0; k < cfg.size(); ++k)
    out << (k ? ";" : "") << cfg[k][0] << ',' << cfg[k][1] << ',' << cfg[k][2];
  out << '\n';
}

}  // namespace

// ---------------------------------------------------------------- boundary conditions
namespace {

const char* const DOF_NAMES[6] = {"X", "Y", "Z", "ThX", "ThY", "ThZ"};  // DOFParameter (Debug form)

void ensure_bc(fem_model* m) {
  const size_t n = size_t(m->nodes_number) * 6;
  if (m->bc_cons.size() != n) {
    m->bc_cons.assign(n, 0);
    m->bc_disp.assign(n, 0.0);
    m->bc_force.assign(n, 0.0);
    m->bc_dirty = true;
  }
}

// kind 0: FEM::add_displacement (methods_for_bc_data_handle.rs:176-203)
// kind 1: FEM::add_concentrated_load (methods_for_bc_data_handle.rs:34-56)
int add_bc(fem_model* m, int kind, size_t n, const uint32_t* numbers, const int32_t* dofs, const double* values,
           size_t* n_ok) {
  if (!m || (n && (!numbers || !dofs || !values))) return FEM_ERR_ARGUMENT;
  size_t i = 0;
  int rc = FEM_OK;
  for (; i < n; ++i) {
    const uint32_t number = numbers[i];
    const int dof = dofs[i];
    if (dof < 0 || dof > 5) {
      rc = fail(m, FEM_ERR_ARGUMENT, "dof must be 0..5 (X, Y, Z, ThX, ThY, ThZ)");
      break;
    }
    if (!node_exists(m, number)) {
      rc = FEM_ERR_VALIDATION;
      break;
    }
    ensure_bc(m);
    const size_t index = size_t(m->node_idx[number]) * 6 + size_t(dof);
    if (kind == 0) {
      if (m->bc_cons[index]) {
        rc = fail(m, FEM_ERR_VALIDATION,
                  std::string("Displacement ") + DOF_NAMES[dof] + " already applied to node " + u(number) + "!");
        break;
      }
      m->bc_cons[index] = 1;
      m->bc_disp[index] = values[i];
    } else {
      m->bc_force[index] += values[i];
    }
    m->bc_dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

int separation_ready(fem_model* m) {
  if (!m) return FEM_ERR_ARGUMENT;
  int64_t rows[4];
  if (!m->dev) return fail(m, FEM_ERR_STATE, "no separation computed: call fem_separate");
  femdev::separation_shape(m->dev, rows, nullptr);
  const std::string e = femdev::separated_vectors(m->dev, nullptr, nullptr, nullptr, nullptr);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_STATE, e + ": call fem_separate");
}

// FEM::add_uniformly_distributed_line_load (methods_for_bc_data_handle.rs:58-100, type 1)
// and add_uniformly_distributed_surface_load (:102-174, type 2): the nodal
// loads are computed on the device and added in call order
int add_distributed(fem_model* m, int type, size_t n, const uint32_t* numbers, const int32_t* dofs,
                    const double* values, size_t* n_ok) {
  if (!m || (n && (!numbers || !dofs || !values))) return FEM_ERR_ARGUMENT;
  static const char* names[3] = {"Truss", "Beam", "Plate"};
  size_t i = 0;
  int rc = FEM_OK;
  std::vector<uint32_t> pos;
  for (; i < n; ++i) {
    if (dofs[i] < 0 || dofs[i] > 5) {
      rc = fail(m, FEM_ERR_ARGUMENT, "dof must be 0..5 (X, Y, Z, ThX, ThY, ThZ)");
      break;
    }
    auto it = m->set[type].by_number.find(numbers[i]);
    if (it == m->set[type].by_number.end()) {
      rc = fail(m, FEM_ERR_VALIDATION, std::string(names[type]) + " element with number " + u(numbers[i]) +
                                           " does not exist!");
      break;
    }
    pos.push_back(it->second);
  }
  if (!pos.empty()) {
    int rc2 = ensure_ready(m);
    if (rc2) {
      if (n_ok) *n_ok = 0;
      return rc2;
    }
    const int per = type == BEAM ? 2 : 4;
    std::vector<double> f(pos.size() * size_t(per));
    auto e = femdev::nodal_loads(m->dev, type, pos.data(), values, pos.size(), f.data());
    if (!e.empty()) {
      if (n_ok) *n_ok = 0;
      return fail(m, FEM_ERR_DEVICE, e);
    }
    ensure_bc(m);
    for (size_t k = 0; k < pos.size(); ++k)
      for (int a = 0; a < per; ++a) {
        const size_t g = size_t(m->set[type].conn[size_t(per) * pos[k] + size_t(a)]) * 6 + size_t(dofs[k]);
        m->bc_force[g] += f[size_t(per) * k + size_t(a)];
      }
    m->bc_dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

int fem_create(double rel_tol, double abs_tol, uint32_t nodes_number, fem_model** out) {
  if (!out) return FEM_ERR_ARGUMENT;
  auto* m = new fem_model();
  m->rel_tol = rel_tol;
  m->abs_tol = abs_tol;
  clear_model(m, nodes_number);
  *out = m;
  return FEM_OK;
}

int fem_reset(fem_model* m, uint32_t nodes_number) {
  if (!m) return FEM_ERR_ARGUMENT;
  clear_model(m, nodes_number);
  return FEM_OK;
}

void fem_destroy(fem_model* m) {
  if (!m) return;
  femdev::stream_destroy(m->side_stream);
  femdev::event_destroy(m->fork_ev);
  femdev::event_destroy(m->join_ev);
  femdev::destroy_device(m->dev);
  delete m;
}

const char* fem_last_error(const fem_model* m) { return m ? m->err.c_str() : "null handle"; }

int fem_add_nodes(fem_model* m, size_t n, const uint32_t* numbers, const double* xyz, size_t* n_ok) {
  if (!m || (n && (!numbers || !xyz))) return FEM_ERR_ARGUMENT;
  size_t i = 0;
  int rc = FEM_OK;
  for (; i < n; ++i) {
    uint32_t number = numbers[i];
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    size_t index = m->node_number.size();
    // methods_for_node_data_handle.rs:39-64
    if (m->nodes_number == 0 || index > size_t(m->nodes_number) - 1) {
      rc = fail(m, FEM_ERR_VALIDATION, "Nodes number could not be greater than " + u(m->nodes_number) + "!");
      break;
    }
    if (m->node_idx.count(number)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Node with number " + u(number) + " already exists!");
      break;
    }
    uint64_t key = coord_key(x, y, z);
    bool dup = false;
    auto it = m->coord_hash.find(key);
    if (it != m->coord_hash.end())
      for (uint32_t o : it->second) {
        const double* c = node_xyz(m, o);
        if (c[0] == x && c[1] == y && c[2] == z) dup = true;
      }
    if (dup) {
      rc = fail(m, FEM_ERR_VALIDATION, "Node with coordinates x: " + rust_debug(x) + ", y: " + rust_debug(y) +
                                           ", z: " + rust_debug(z) + " already exists!");
      break;
    }
    m->node_idx[number] = uint32_t(index);
    m->node_number.push_back(number);
    m->xyz.insert(m->xyz.end(), {x, y, z});
    m->coord_hash[key].push_back(uint32_t(index));
    m->node_plates.emplace_back();
    m->dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

int fem_add_trusses(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* E,
                    const double* A, const double* A2, size_t* n_ok) {
  if (!m || (n && (!numbers || !conn || !E || !A))) return FEM_ERR_ARGUMENT;
  size_t i = 0;
  int rc = FEM_OK;
  auto& S = m->set[TRUSS];
  for (; i < n; ++i) {
    uint32_t number = numbers[i], n1 = conn[2 * i], n2 = conn[2 * i + 1];
    double a2 = A2 ? A2[i] : NAN;
    bool has_a2 = !std::isnan(a2);
    // methods_for_truss_data_handle.rs:49-67, truss.rs:44-64
    if (!node_exists(m, n1) || !node_exists(m, n2)) { rc = FEM_ERR_VALIDATION; break; }
    if (S.by_number.count(number)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Truss element with number " + u(number) + " already exists!");
      break;
    }
    if (m->truss_pairs.count(pair_key(n1, n2))) {
      rc = fail(m, FEM_ERR_VALIDATION, "Truss element with node number " + u(n1) + " and " + u(n2) + " already exists!");
      break;
    }
    if (!positive_or(m, E[i], "Young's modulus", E[i]) || !positive_or(m, A[i], "Area", A[i]) ||
        (has_a2 && !positive_or(m, a2, "Area2", a2))) {
      rc = FEM_ERR_VALIDATION;
      break;
    }
    uint32_t i1 = m->node_idx[n1], i2 = m->node_idx[n2];
    const double *p1 = node_xyz(m, i1), *p2 = node_xyz(m, i2);
    if (p1[0] == p2[0] && p1[1] == p2[1] && p1[2] == p2[2]) {
      rc = fail(m, FEM_ERR_VALIDATION, "Zero length vector could not be aligned!");
      break;
    }
    S.by_number[number] = uint32_t(S.size());
    S.number.push_back(number);
    S.conn.insert(S.conn.end(), {i1, i2});
    S.params.insert(S.params.end(), {E[i], A[i], has_a2 ? a2 : NAN});
    S.seq.push_back(m->seq++);
    m->truss_pairs[pair_key(n1, n2)] = number;
    m->dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

int fem_add_beams(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* props,
                  size_t* n_ok) {
  if (!m || (n && (!numbers || !conn || !props))) return FEM_ERR_ARGUMENT;
  size_t i = 0;
  int rc = FEM_OK;
  auto& S = m->set[BEAM];
  for (; i < n; ++i) {
    uint32_t number = numbers[i], n1 = conn[2 * i], n2 = conn[2 * i + 1];
    const double* p = props + 11 * i;
    // methods_for_beam_data_handle.rs:49-68, beam.rs:63-117
    if (!node_exists(m, n1) || !node_exists(m, n2)) { rc = FEM_ERR_VALIDATION; break; }
    if (S.by_number.count(number)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Beam element with number " + u(number) + " already exists!");
      break;
    }
    if (m->beam_pairs.count(pair_key(n1, n2))) {
      rc = fail(m, FEM_ERR_VALIDATION, "Beam element with node number " + u(n1) + " and " + u(n2) + " already exists!");
      break;
    }
    if (!positive_or(m, p[0], "Young's modulus", p[0]) || !positive_or(m, p[1], "Poisson's ratio", p[0]) ||
        !positive_or(m, p[2], "Area", p[2]) || !positive_or(m, p[3], "I11", p[3]) ||
        !positive_or(m, p[4], "I22", p[4]) || !positive_or(m, p[6], "It", p[6]) ||
        !positive_or(m, p[7], "Shear factor", p[7])) {
      rc = FEM_ERR_VALIDATION;
      break;
    }
    uint32_t i1 = m->node_idx[n1], i2 = m->node_idx[n2];
    const double *a = node_xyz(m, i1), *b = node_xyz(m, i2);
    double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double f = ((p[8] * d[0] + p[9] * d[1]) + p[10] * d[2]) / ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    double pp[3] = {p[8] - f * d[0], p[9] - f * d[1], p[10] - f * d[2]};
    if (fem::norm3(pp[0], pp[1], pp[2]) == 0.0) {
      rc = fail(m, FEM_ERR_VALIDATION, "Local axis 1 direction [" + rust_debug(p[8]) + ", " + rust_debug(p[9]) + ", " +
                                           rust_debug(p[10]) + "] parallel to element " + u(number) + "!");
      break;
    }
    if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) {
      rc = fail(m, FEM_ERR_VALIDATION, "Zero length vector could not be aligned!");
      break;
    }
    S.by_number[number] = uint32_t(S.size());
    S.number.push_back(number);
    S.conn.insert(S.conn.end(), {i1, i2});
    S.params.insert(S.params.end(), p, p + 11);
    S.seq.push_back(m->seq++);
    m->beam_pairs[pair_key(n1, n2)] = number;
    m->dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

int fem_add_plates(fem_model* m, size_t n, const uint32_t* numbers, const uint32_t* conn, const double* props,
                   size_t* n_ok) {
  if (!m || (n && (!numbers || !conn || !props))) return FEM_ERR_ARGUMENT;
  size_t i = 0;
  int rc = FEM_OK;
  auto& S = m->set[PLATE];
  for (; i < n; ++i) {
    uint32_t number = numbers[i];
    const uint32_t* nn = conn + 4 * i;
    const double* p = props + 4 * i;
    // methods_for_plate_data_handle.rs:62-102, plate.rs:58-158
    bool ok = true;
    for (int k = 0; k < 4 && ok; ++k) ok = node_exists(m, nn[k]);
    if (!ok) { rc = FEM_ERR_VALIDATION; break; }
    if (S.by_number.count(number)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Plate element with number " + u(number) + " already exists!");
      break;
    }
    uint32_t idx[4];
    for (int k = 0; k < 4; ++k) idx[k] = m->node_idx[nn[k]];
    bool same = false;
    for (uint32_t cand : m->node_plates[idx[0]]) {
      const uint32_t* c = &S.conn[4 * size_t(cand)];
      bool all = true;
      for (int k = 0; k < 4 && all; ++k) all = idx[k] == c[0] || idx[k] == c[1] || idx[k] == c[2] || idx[k] == c[3];
      if (all) { same = true; break; }
    }
    if (same) {
      rc = fail(m, FEM_ERR_VALIDATION, "Plate element with nodes numbers [" + u(nn[0]) + ", " + u(nn[1]) + ", " +
                                           u(nn[2]) + ", " + u(nn[3]) + "] already exists!");
      break;
    }
    if (!positive_or(m, p[0], "Young's modulus", p[0]) || !positive_or(m, p[1], "Poisson's ratio", p[0]) ||
        !positive_or(m, p[2], "Thickness", p[0]) || !positive_or(m, p[3], "Shear factor", p[3])) {
      rc = FEM_ERR_VALIDATION;
      break;
    }
    const double* pts[4] = {node_xyz(m, idx[0]), node_xyz(m, idx[1]), node_xyz(m, idx[2]), node_xyz(m, idx[3])};
    if (quad_on_line(pts)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Some nodes of " + u(number) + " element lie on the line!");
      break;
    }
    if (!quad_on_plane(pts, m->abs_tol)) {
      rc = fail(m, FEM_ERR_VALIDATION, "Not all nodes of element " + u(number) + " lie on the plane!");
      break;
    }
    int hull = quad_hull_points(nn, pts, m->abs_tol);
    if (hull < 0) {
      rc = fail(m, FEM_ERR_VALIDATION, "Zero length vector could not be aligned!");
      break;
    }
    if (hull != 4) {
      rc = fail(m, FEM_ERR_VALIDATION, "Element " + u(number) + " non-convex!");
      break;
    }
    uint32_t pos = uint32_t(S.size());
    S.by_number[number] = pos;
    S.number.push_back(number);
    S.conn.insert(S.conn.end(), idx, idx + 4);
    S.params.insert(S.params.end(), p, p + 4);
    S.seq.push_back(m->seq++);
    for (int k = 0; k < 4; ++k)
      if (k == 0 || std::find(idx, idx + k, idx[k]) == idx + k) m->node_plates[idx[k]].push_back(pos);
    m->dirty = true;
  }
  if (n_ok) *n_ok = i;
  return rc;
}

int fem_get_rotation(fem_model* m, int type, uint32_t number, double out[9]) {
  if (!m || !out || type < 0 || type > 2) return FEM_ERR_ARGUMENT;
  static const char* names[3] = {"Truss", "Beam", "Plate"};
  auto it = m->set[type].by_number.find(number);
  if (it == m->set[type].by_number.end())
    return fail(m, FEM_ERR_VALIDATION, std::string(names[type]) + " element with number " + u(number) + " does not exist!");
  int rc = ensure_ready(m);
  if (rc) return rc;
  auto e = femdev::get_rotation(m->dev, type, it->second, out);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

int fem_build_pattern(fem_model* m) {
  if (!m) return FEM_ERR_ARGUMENT;
  return ensure_ready(m);
}

int fem_quad_points_on_same_line(const double* points, int* out) {
  if (!points || !out) return FEM_ERR_ARGUMENT;
  const double* p[4] = {points, points + 3, points + 6, points + 9};
  *out = quad_on_line(p) ? 1 : 0;
  return FEM_OK;
}

int fem_quad_points_on_same_plane(const double* points, double abs_tol, int* out) {
  if (!points || !out) return FEM_ERR_ARGUMENT;
  const double* p[4] = {points, points + 3, points + 6, points + 9};
  *out = quad_on_plane(p, abs_tol) ? 1 : 0;
  return FEM_OK;
}

int fem_convex_hull_on_four_points_on_plane(const uint32_t* numbers, const double* points, double rel_tol,
                                            double abs_tol, uint32_t* hull, size_t* n_hull, char* err,
                                            size_t err_cap) {
  (void)rel_tol;  // the restated alignment snaps with abs_tol only
  if (!numbers || !points || !hull || !n_hull) return FEM_ERR_ARGUMENT;
  const double* p[4] = {points, points + 3, points + 6, points + 9};
  std::vector<uint32_t> h;
  if (!quad_hull(numbers, p, abs_tol, h)) {
    static const char msg[] = "Zero length vector could not be aligned!";
    if (err && err_cap) {
      std::strncpy(err, msg, err_cap - 1);
      err[err_cap - 1] = '\0';
    }
    *n_hull = 0;
    return FEM_ERR_VALIDATION;
  }
  for (size_t i = 0; i < h.size(); ++i) hull[i] = h[i];
  *n_hull = h.size();
  return FEM_OK;
}

int fem_assemble_into(fem_model* m, double* val_device, void* stream) {
  if (!m || (!val_device && m->nnz)) return FEM_ERR_ARGUMENT;
  int rc = ensure_ready(m);
  if (rc) return rc;
  const bool t = m->time_assembly;
  std::string e = m->path == "plate_fused" ? femdev::assemble_plate_fused(m->dev, val_device, stream, t)
                  : m->path == "plate_lanes" ? femdev::assemble_plate_lanes(m->dev, val_device, stream, t)
                  : m->path == "plate_blocks" ? femdev::assemble_plate_blocks(m->dev, val_device, stream, t)
                  : m->path == "plate_corners" ? femdev::assemble_plate_corners(m->dev, val_device, stream, t)
                  : m->path == "bar_owner" ? femdev::assemble_bar_owner(m->dev, val_device, stream, t)
                                           : femdev::assemble_generic(m->dev, val_device, stream, t);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  if (val_device == femdev::values(m->dev)) m->own_vals = true;
  return FEM_OK;
}

int fem_assemble(fem_model* m, void* stream) {
  if (!m) return FEM_ERR_ARGUMENT;
  int rc = ensure_ready(m);
  if (rc) return rc;
  return fem_assemble_into(m, femdev::values(m->dev), stream);
}

int fem_csr_info(fem_model* m, int64_t* n_rows, int64_t* nnz) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->pattern_valid || m->dirty) return fail(m, FEM_ERR_STATE, "the pattern is not built: call fem_assemble");
  if (n_rows) *n_rows = int64_t(m->nodes_number) * 6;
  if (nnz) *nnz = m->nnz;
  return FEM_OK;
}

int fem_csr_device(fem_model* m, const int64_t** rp, const int32_t** ci, const double** v) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->pattern_valid || m->dirty) return fail(m, FEM_ERR_STATE, "the pattern is not built: call fem_assemble");
  if (rp) *rp = femdev::row_ptr(m->dev);
  if (ci) *ci = femdev::col_idx(m->dev);
  if (v) *v = femdev::values(m->dev);
  return FEM_OK;
}

int fem_csr_to_host(fem_model* m, int64_t* rp, int32_t* ci, double* v) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->pattern_valid || m->dirty) return fail(m, FEM_ERR_STATE, "the pattern is not built: call fem_assemble");
  auto e = femdev::copy_csr_to_host(m->dev, int64_t(m->nodes_number) * 6, m->nnz, rp, ci, v);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

double fem_last_assembly_ms(fem_model* m) { return m && m->dev ? femdev::last_ms(m->dev) : -1.0; }

int fem_backward(fem_model* m, const double* dl_dk, double* dl_dparam_truss, double* dl_dparam_beam,
                 double* dl_dparam_plate, double* dl_dxyz, void* stream) {
  return run_backward(m, 0, dl_dk, dl_dparam_truss, dl_dparam_beam, dl_dparam_plate, dl_dxyz, stream);
}

int fem_compliance_sensitivity(fem_model* m, const double* u, double* s_truss, double* s_beam, double* s_plate,
                               double* s_xyz, void* stream) {
  return run_backward(m, 1, u, s_truss, s_beam, s_plate, s_xyz, stream);
}

int fem_forward_backward(fem_model* m, double* val_device, const double* dl_dk, double* dl_dparam_truss,
                         double* dl_dparam_beam, double* dl_dparam_plate, double* dl_dxyz, void* stream) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->side_stream) {
    std::string e = femdev::stream_create(m->side_stream);
    if (e.empty()) e = femdev::event_create(m->fork_ev);
    if (e.empty()) e = femdev::event_create(m->join_ev);
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  }
  int rc = ensure_adjoint(m);
  if (rc) return rc;
  // the backward does not read the forward's values: both passes run at once
  auto e = femdev::stream_wait(m->side_stream, stream, m->fork_ev);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  rc = run_backward(m, 0, dl_dk, dl_dparam_truss, dl_dparam_beam, dl_dparam_plate, dl_dxyz, m->side_stream);
  if (rc) return rc;
  rc = fem_assemble_into(m, val_device, stream);
  if (rc) return rc;
  e = femdev::stream_wait(stream, m->side_stream, m->join_ev);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

int fem_model_size(const fem_model* m, size_t* n_nodes, size_t* n_truss, size_t* n_beam, size_t* n_plate) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (n_nodes) *n_nodes = m->node_number.size();
  if (n_truss) *n_truss = m->set[TRUSS].size();
  if (n_beam) *n_beam = m->set[BEAM].size();
  if (n_plate) *n_plate = m->set[PLATE].size();
  return FEM_OK;
}

double fem_last_backward_ms(fem_model* m) { return m && m->dev ? femdev::last_backward_ms(m->dev) : -1.0; }

int fem_update_params(fem_model* m, int type, const double* params_device, void* stream) {
  if (!m || type < 0 || type > 2 || (!params_device && m->set[type].size())) return FEM_ERR_ARGUMENT;
  int rc = ensure_ready(m);
  if (rc) return rc;
  auto e = femdev::replace_params(m->dev, type, params_device, m->set[type].params.data(), stream);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  m->own_vals = false;
  if (type == PLATE && m->path == "plate_lanes") {  // the slot layout follows the new parameters
    double mat[3];
    e = femdev::set_plate_lane_material(m->dev, plate_material_uniform(m, mat), mat);
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  }
  return type == BEAM ? rerun_setup(m, false, BEAM) : FEM_OK;
}

int fem_update_nodes(fem_model* m, const double* xyz_device, void* stream) {
  if (!m || (!xyz_device && m->node_number.size())) return FEM_ERR_ARGUMENT;
  int rc = ensure_ready(m);
  if (rc) return rc;
  auto e = femdev::replace_nodes(m->dev, xyz_device, m->xyz.data(), stream);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  m->own_vals = false;
  m->coord_hash.clear();
  for (size_t i = 0; i < m->node_number.size(); ++i)
    m->coord_hash[coord_key(m->xyz[3 * i], m->xyz[3 * i + 1], m->xyz[3 * i + 2])].push_back(uint32_t(i));
  return rerun_setup(m, true, 0);
}

int fem_add_displacements(fem_model* m, size_t n, const uint32_t* numbers, const int32_t* dofs, const double* values,
                          size_t* n_ok) {
  return add_bc(m, 0, n, numbers, dofs, values, n_ok);
}

int fem_add_concentrated_loads(fem_model* m, size_t n, const uint32_t* numbers, const int32_t* dofs,
                               const double* values, size_t* n_ok) {
  return add_bc(m, 1, n, numbers, dofs, values, n_ok);
}

int fem_add_uniformly_distributed_line_loads(fem_model* m, size_t n, const uint32_t* beam_numbers,
                                             const int32_t* dofs, const double* values, size_t* n_ok) {
  return add_distributed(m, BEAM, n, beam_numbers, dofs, values, n_ok);
}

int fem_add_uniformly_distributed_surface_loads(fem_model* m, size_t n, const uint32_t* plate_numbers,
                                                const int32_t* dofs, const double* values, size_t* n_ok) {
  return add_distributed(m, PLATE, n, plate_numbers, dofs, values, n_ok);
}

int fem_separate(fem_model* m, void* stream, int64_t* n_aa, int64_t* n_bb, int64_t nnz[4]) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->pattern_valid || m->dirty || !m->own_vals || !m->dev)
    return fail(m, FEM_ERR_STATE, "the stiffness is not assembled: call fem_assemble");
  ensure_bc(m);
  if (m->bc_dirty) {
    auto e = femdev::upload_bc(m->dev, m->bc_cons.data(), m->bc_disp.data(), m->bc_force.data(), m->bc_cons.size());
    if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
    m->bc_dirty = false;
  }
  femdev::SepResult r;
  auto e = femdev::separate_sparse(m->dev, r, stream, true, m->sep_lanes);
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  // methods_for_separate_stiffness_matrix.rs:232-259, 304-308
  if (r.err_row >= 0) {
    const size_t node = size_t(r.err_row) / 6;
    const uint32_t number = node < m->node_number.size() ? m->node_number[node] : 0;
    femdev::invalidate_separation(m->dev);
    return fail(m, FEM_ERR_VALIDATION, std::string("There are no stiffness to withstand displacement ") +
                                           DOF_NAMES[r.err_row % 6] + " applied to node " + u(number) + "!");
  }
  if (r.n_bb == 0) {
    femdev::invalidate_separation(m->dev);
    return fail(m, FEM_ERR_VALIDATION, "No restraints");
  }
  if (r.nnz[0] == 0) {
    femdev::invalidate_separation(m->dev);
    return fail(m, FEM_ERR_VALIDATION, "Sparse separation: K_aa is empty (structure has no free stiffness?)");
  }
  if (n_aa) *n_aa = r.n_aa;
  if (n_bb) *n_bb = r.n_bb;
  if (nnz)
    for (int w = 0; w < 4; ++w) nnz[w] = r.nnz[w];
  return FEM_OK;
}

int fem_separated_device(fem_model* m, int which, const int64_t** row_ptr, const int32_t** col_idx,
                         const double** values) {
  int rc = separation_ready(m);
  if (rc) return rc;
  if (which < 0 || which > 3) return fail(m, FEM_ERR_ARGUMENT, "which must be 0..3 (aa, ab, ba, bb)");
  auto e = femdev::separated_csr(m->dev, which, row_ptr, col_idx, values);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_STATE, e);
}

int fem_separated_to_host(fem_model* m, int which, int64_t* row_ptr, int32_t* col_idx, double* values) {
  int rc = separation_ready(m);
  if (rc) return rc;
  if (which < 0 || which > 3) return fail(m, FEM_ERR_ARGUMENT, "which must be 0..3 (aa, ab, ba, bb)");
  int64_t rows[4], nnz[4];
  femdev::separation_shape(m->dev, rows, nnz);
  const int64_t* rp = nullptr;
  const int32_t* ci = nullptr;
  const double* v = nullptr;
  auto e = femdev::separated_csr(m->dev, which, &rp, &ci, &v);
  if (e.empty() && row_ptr) e = femdev::copy_to_host(row_ptr, rp, sizeof(int64_t) * size_t(rows[which] + 1));
  if (e.empty() && col_idx && nnz[which]) e = femdev::copy_to_host(col_idx, ci, sizeof(int32_t) * size_t(nnz[which]));
  if (e.empty() && values && nnz[which]) e = femdev::copy_to_host(values, v, sizeof(double) * size_t(nnz[which]));
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

int fem_separated_vectors_device(fem_model* m, const int64_t** k_aa_indexes, const int64_t** k_bb_indexes,
                                 const double** r_a, const double** u_b) {
  int rc = separation_ready(m);
  if (rc) return rc;
  auto e = femdev::separated_vectors(m->dev, k_aa_indexes, k_bb_indexes, r_a, u_b);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_STATE, e);
}

int fem_separated_vectors_to_host(fem_model* m, int64_t* k_aa_indexes, int64_t* k_bb_indexes, double* r_a,
                                  double* u_b) {
  int rc = separation_ready(m);
  if (rc) return rc;
  int64_t rows[4];
  femdev::separation_shape(m->dev, rows, nullptr);
  const int64_t *aa = nullptr, *bb = nullptr;
  const double *ra = nullptr, *ub = nullptr;
  auto e = femdev::separated_vectors(m->dev, &aa, &bb, &ra, &ub);
  const size_t na = size_t(rows[0]), nb = size_t(rows[2]);
  if (e.empty() && k_aa_indexes && na) e = femdev::copy_to_host(k_aa_indexes, aa, sizeof(int64_t) * na);
  if (e.empty() && k_bb_indexes && nb) e = femdev::copy_to_host(k_bb_indexes, bb, sizeof(int64_t) * nb);
  if (e.empty() && r_a && na) e = femdev::copy_to_host(r_a, ra, sizeof(double) * na);
  if (e.empty() && u_b && nb) e = femdev::copy_to_host(u_b, ub, sizeof(double) * nb);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

double fem_last_separate_ms(fem_model* m) { return m && m->dev ? femdev::last_separate_ms(m->dev) : -1.0; }

int fem_solve_pcg(fem_model* m, int preconditioner, uint64_t max_iter, void* stream, uint64_t* iterations,
                  int* converged) {
  int rc = separation_ready(m);
  if (rc) return rc;
  if (preconditioner != 0 && preconditioner != 1)
    return fail(m, FEM_ERR_ARGUMENT, "preconditioner must be 0 (Jacobi) or 1 (node-block Jacobi)");
  if (max_iter > uint64_t(INT32_MAX)) return fail(m, FEM_ERR_ARGUMENT, "max_iter too large");
  femdev::SolveResult r;
  auto e = femdev::solve_pcg(m->dev, preconditioner, int(max_iter), m->rel_tol, m->abs_tol, stream, r, m->wide_loads);
  if (!e.empty()) {
    const bool numeric = e.rfind("PCG", 0) == 0;
    return fail(m, numeric ? FEM_ERR_VALIDATION : FEM_ERR_DEVICE, e);
  }
  if (iterations) *iterations = uint64_t(r.iterations);
  if (converged) *converged = r.converged ? 1 : 0;
  return FEM_OK;
}

int fem_solve_pcg_noalloc(fem_model* m, int preconditioner, uint64_t max_iter, double* u_a, double* r_r,
                          void* stream, uint64_t* iterations, int* converged) {
  int rc = separation_ready(m);
  if (rc) return rc;
  if (preconditioner != 0 && preconditioner != 1)
    return fail(m, FEM_ERR_ARGUMENT, "preconditioner must be 0 (Jacobi) or 1 (node-block Jacobi)");
  if (max_iter > uint64_t(INT32_MAX)) return fail(m, FEM_ERR_ARGUMENT, "max_iter too large");
  femdev::SolveResult r;
  auto e = femdev::solve_pcg_static(m->dev, preconditioner, int(max_iter), m->rel_tol, m->abs_tol, stream, u_a, r_r,
                                    r, m->wide_loads);
  if (!e.empty()) {
    const bool numeric = e.rfind("PCG", 0) == 0;
    const bool limit = e.find("static scratch") != std::string::npos || e.find("device buffers") != std::string::npos;
    return fail(m, numeric ? FEM_ERR_VALIDATION : limit ? FEM_ERR_ARGUMENT : FEM_ERR_DEVICE, e);
  }
  if (iterations) *iterations = uint64_t(r.iterations);
  if (converged) *converged = r.converged ? 1 : 0;
  return FEM_OK;
}

int fem_solution_device(fem_model* m, const double** u_a, const double** r_r) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->dev) return fail(m, FEM_ERR_STATE, "no solution computed: call fem_solve_pcg");
  auto e = femdev::solution(m->dev, u_a, r_r, nullptr, nullptr);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_STATE, e + ": call fem_solve_pcg");
}

int fem_solution_to_host(fem_model* m, double* u_a, double* r_r) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->dev) return fail(m, FEM_ERR_STATE, "no solution computed: call fem_solve_pcg");
  const double *ua = nullptr, *rr = nullptr;
  int64_t na = 0, nb = 0;
  auto e = femdev::solution(m->dev, &ua, &rr, &na, &nb);
  if (!e.empty()) return fail(m, FEM_ERR_STATE, e + ": call fem_solve_pcg");
  if (u_a && na) e = femdev::copy_to_host(u_a, ua, sizeof(double) * size_t(na));
  if (e.empty() && r_r && nb) e = femdev::copy_to_host(r_r, rr, sizeof(double) * size_t(nb));
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

double fem_last_solve_ms(fem_model* m) { return m && m->dev ? femdev::last_solve_ms(m->dev) : -1.0; }

int fem_compose_global_result(fem_model* m) {
  if (!m) return FEM_ERR_ARGUMENT;
  if (!m->dev) return fail(m, FEM_ERR_STATE, "no solution computed: call fem_solve_pcg");
  const double *ua = nullptr, *rr = nullptr;
  int64_t na = 0, nb = 0;
  auto e = femdev::solution(m->dev, &ua, &rr, &na, &nb);
  if (!e.empty()) return fail(m, FEM_ERR_STATE, e + ": call fem_solve_pcg");
  const int64_t *aa = nullptr, *bb = nullptr;
  e = femdev::separated_vectors(m->dev, &aa, &bb, nullptr, nullptr);
  if (!e.empty()) return fail(m, FEM_ERR_STATE, e);
  std::vector<int64_t> aah(static_cast<size_t>(na)), bbh(static_cast<size_t>(nb));
  std::vector<double> uah(static_cast<size_t>(na)), rrh(static_cast<size_t>(nb));
  if (na) e = femdev::copy_to_host(aah.data(), aa, sizeof(int64_t) * size_t(na));
  if (e.empty() && nb) e = femdev::copy_to_host(bbh.data(), bb, sizeof(int64_t) * size_t(nb));
  if (e.empty() && na) e = femdev::copy_to_host(uah.data(), ua, sizeof(double) * size_t(na));
  if (e.empty() && nb) e = femdev::copy_to_host(rrh.data(), rr, sizeof(double) * size_t(nb));
  if (!e.empty()) return fail(m, FEM_ERR_DEVICE, e);
  ensure_bc(m);
  // methods_for_global_analysis.rs:335-358
  for (int64_t i = 0; i < na; ++i) m->bc_disp[size_t(aah[size_t(i)])] = uah[size_t(i)];
  for (int64_t i = 0; i < nb; ++i) m->bc_force[size_t(bbh[size_t(i)])] = rrh[size_t(i)];
  m->bc_dirty = true;
  return FEM_OK;
}

int fem_solution_displacements(fem_model* m, void* stream, const double** u) {
  if (!m || !u) return FEM_ERR_ARGUMENT;
  if (!m->dev) return fail(m, FEM_ERR_STATE, "no solution computed: call fem_solve_pcg");
  auto e = femdev::solution_displacements(m->dev, stream, u);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_STATE, e + ": call fem_solve_pcg");
}

int fem_element_forces(fem_model* m, const double* u_device, double* truss_out, double* beam_out, double* plate_out,
                       void* stream) {
  if (!m || !u_device) return FEM_ERR_ARGUMENT;
  int rc = ensure_ready(m);
  if (rc) return rc;
  double* const out[3] = {truss_out, beam_out, plate_out};
  auto e = femdev::element_forces(m->dev, u_device, out, stream);
  return e.empty() ? FEM_OK : fail(m, FEM_ERR_DEVICE, e);
}

int fem_global_result(fem_model* m, uint32_t* node_numbers, int32_t* dofs, double* displacements, double* loads) {
  if (!m) return FEM_ERR_ARGUMENT;
  ensure_bc(m);
  // methods_for_global_analysis.rs:360-384, nodes in insertion order
  const size_t nn = m->node_number.size();
  for (size_t i = 0; i < nn; ++i)
    for (int k = 0; k < 6; ++k) {
      const size_t o = 6 * i + size_t(k), g = 6 * size_t(i) + size_t(k);
      if (node_numbers) node_numbers[o] = m->node_number[i];
      if (dofs) dofs[o] = k;
      if (displacements) displacements[o] = m->bc_disp[g];
      if (loads) loads[o] = m->bc_force[g];
    }
  return FEM_OK;
}

const char* fem_assembly_path(fem_model* m) { return m ? m->path.c_str() : ""; }

int fem_set_option(fem_model* m, const char* key, int64_t value) {
  if (!m || !key) return FEM_ERR_ARGUMENT;
  std::string k(key);
  if (k == "plate_patch_nodes") {
    if (value < 4 || value > femdev::PF_MAX_NODES)
      return fail(m, FEM_ERR_ARGUMENT, "plate_patch_nodes must be in [4, 64]");
    m->pf_nodes = int(value);
  } else if (k == "plate_wg_per_cu") {
    if (value < 0 || value > 64) return fail(m, FEM_ERR_ARGUMENT, "plate_wg_per_cu must be in [0, 64]");
    m->pf_wg_per_cu = int(value);
  } else if (k == "force_generic") {
    m->force_generic = value != 0;
  } else if (k == "plate_kernel") {
    if (value < 0 || value > 3) return fail(m, FEM_ERR_ARGUMENT, "plate_kernel must be 0, 1, 2 or 3");
    m->pf_kernel = int(value);
  } else if (k == "plate_patch_order") {
    if (value < 0 || value > 2) return fail(m, FEM_ERR_ARGUMENT, "plate_patch_order must be 0, 1 or 2");
    m->pl_order = int(value);
  } else if (k == "separate_lanes") {
    if (value != 0 && value != 4 && value != 8 && value != 16 && value != 32)
      return fail(m, FEM_ERR_ARGUMENT, "separate_lanes must be 0 (automatic), 4, 8, 16 or 32");
    m->sep_lanes = int(value);
    return FEM_OK;  // speed only; the pattern stays
  } else if (k == "wide_loads") {
    m->wide_loads = value != 0;
    return FEM_OK;  // speed only
  } else if (k == "time_assembly") {
    // 0: no events around the assembly launches (two fewer packets between
    // back-to-back assemblies); fem_last_assembly_ms then keeps the last timed one
    m->time_assembly = value != 0;
    return FEM_OK;
  } else {
    return fail(m, FEM_ERR_ARGUMENT, "unknown option " + k);
  }
  m->pattern_valid = false;
  m->own_vals = false;
  return FEM_OK;
}

int fem_batch_create(fem_model* const* models, size_t n, fem_batch** out) {
  if (!out || (!models && n)) return FEM_ERR_ARGUMENT;
  *out = nullptr;
  for (size_t i = 0; i < n; ++i)
    if (!models[i] || models[i]->rel_tol != models[0]->rel_tol || models[i]->abs_tol != models[0]->abs_tol)
      return FEM_ERR_ARGUMENT;
  auto* b = new fem_batch();
  // Groups, so that every mesh takes the same value path as when assembled
  // alone: flat plate-only meshes (every frame the identity: fused plate
  // kernel), split so that each group's values stay below the fused kernel's
  // 32-bit byte offsets (at most 118 values per node at valence <= 8); meshes
  // of bars with beams (node-owner kernel); truss-only meshes (the same
  // kernels' truss ranges, in a group of their own so that its launch runs
  // beside the beam group's); all other meshes (generic path).
  const uint64_t plate_cap = (uint64_t(1) << 32) / 8 / 118;  // nodes per plate group
  std::vector<int> grp(n, -1);
  std::vector<uint64_t> gnodes;
  int cur_plate = -1, bars = -1, trusses = -1, generic = -1;
  auto open_group = [&](int& g) {
    if (g < 0) {
      g = int(gnodes.size());
      gnodes.push_back(0);
    }
    return g;
  };
  for (size_t i = 0; i < n; ++i) {
    const fem_model* m = models[i];
    const uint64_t nn = m->node_number.size();
    const bool plates_only = m->set[PLATE].size() && !m->set[TRUSS].size() && !m->set[BEAM].size();
    if (plates_only && nn <= plate_cap && plates_flat_identity(m)) {
      if (cur_plate >= 0 && gnodes[size_t(cur_plate)] + nn > plate_cap) cur_plate = -1;
      grp[i] = open_group(cur_plate);
    } else if (!m->set[PLATE].size() && !m->set[BEAM].size()) {
      grp[i] = open_group(trusses);
    } else if (!m->set[PLATE].size()) {
      grp[i] = open_group(bars);
    } else {
      grp[i] = open_group(generic);
    }
    gnodes[size_t(grp[i])] += nn;
  }
  for (uint64_t v : gnodes)
    if (v >= (uint64_t(1) << 31)) {
      delete b;
      return FEM_ERR_ARGUMENT;
    }
  for (size_t k = 0; k < gnodes.size(); ++k) {
    auto* g = new fem_model();
    g->rel_tol = models[0]->rel_tol;
    g->abs_tol = models[0]->abs_tol;
    g->nodes_number = uint32_t(gnodes[k]);
    b->groups.push_back(g);
  }
  for (size_t i = 0; i < n; ++i) {
    fem_model* g = b->groups[size_t(grp[i])];
    fem_batch::Mesh mm;
    mm.group = grp[i];
    mm.node_off = g->node_number.size();
    mm.n_nodes = models[i]->node_number.size();
    b->meshes.push_back(mm);
    append_model(g, models[i], uint32_t(i));
  }
  auto e = femdev::stream_create(b->stream);
  b->timer = femdev::timer_create();
  if (!e.empty() || !b->timer) {
    fem_batch_destroy(b);
    return FEM_ERR_DEVICE;
  }
  *out = b;
  return FEM_OK;
}

void fem_batch_destroy(fem_batch* b) {
  if (!b) return;
  femdev::graph_destroy(b->graph);
  femdev::timer_destroy(b->timer);
  for (auto* g : b->groups) fem_destroy(g);
  for (void* s : b->side) femdev::stream_destroy(s);
  for (void* ev : b->fork_ev) femdev::event_destroy(ev);
  for (void* ev : b->join_ev) femdev::event_destroy(ev);
  femdev::stream_destroy(b->stream);
  delete b;
}

const char* fem_batch_last_error(const fem_batch* b) { return b ? b->err.c_str() : "null handle"; }

int fem_batch_size(const fem_batch* b, size_t* n_meshes, size_t* n_groups) {
  if (!b) return FEM_ERR_ARGUMENT;
  if (n_meshes) *n_meshes = b->meshes.size();
  if (n_groups) *n_groups = b->groups.size();
  return FEM_OK;
}

int fem_batch_autotune(fem_batch* b) {
  if (!b) return FEM_ERR_ARGUMENT;
  // tuned shapes are cached per (gfx arch, batch signature): for the life of
  // the process and in the tune cache file
  static std::map<std::string, std::pair<int, std::vector<std::array<int, 3>>>> cache;
  std::string sig;
  for (auto* g : b->groups)
    sig += std::to_string(g->set[0].size()) + "," + std::to_string(g->set[1].size()) + "," +
           std::to_string(g->set[2].size()) + "," + std::to_string(g->node_number.size()) + ";";
  sig += std::to_string(b->meshes.size());
  const std::string key = femdev::device_arch() + " " + sig;
  std::vector<std::array<int, 3>> best;  // per group: patch nodes, workgroups per CU, generic block
  int best_schedule = -1;                // -1: to be timed
  auto it = cache.find(key);
  if (it != cache.end()) {
    best_schedule = it->second.first;
    best = it->second.second;
  } else if (tune_cache_find(key, best_schedule, best) && best.size() == b->groups.size()) {
    cache[key] = {best_schedule, best};
  } else {
    best.clear();
    best_schedule = -1;
    for (auto* g : b->groups) {
      std::array<int, 3> cfg = {g->pf_nodes, g->pf_wg_per_cu, 256};
      double best_ms = 1e300;
      const bool plate_only = g->set[PLATE].size() && !g->set[TRUSS].size() && !g->set[BEAM].size();
      const int patch_opts[2] = {32, 64};
      for (int pn : patch_opts) {
        // patch_nodes shapes only the row-staging plate kernel
        if ((!plate_only || g->pf_kernel >= 1) && pn != 64) continue;
        if (g->pf_nodes != pn) {
          g->pf_nodes = pn;
          g->pattern_valid = false;
          b->ready = false;
        }
        int rc = batch_ready(b);
        if (rc) return rc;
        if (g->path == "plate_fused") {
          for (int wg : {1, 2, 0}) {
            auto e = femdev::set_plate_grid(g->dev, wg);
            if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
            double ms = 0;
            rc = time_group(b, g, 3, ms);
            if (rc) return rc;
            if (ms < best_ms) { best_ms = ms; cfg = {pn, wg, 256}; }
          }
        } else if (g->path == "plate_blocks" || g->path == "plate_corners") {
          for (int wg : {0, 4}) {
            auto e = femdev::set_plate_block_grid(g->dev, wg);
            if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
            double ms = 0;
            rc = time_group(b, g, 3, ms);
            if (rc) return rc;
            if (ms < best_ms) { best_ms = ms; cfg = {pn, wg, 256}; }
          }
        } else if (g->path == "bar_owner" || g->path == "plate_lanes") {
          cfg = {pn, g->pf_wg_per_cu, 256};  // fixed workgroup shape
        } else {
          for (int bs : {128, 256, 512}) {
            femdev::set_generic_block(g->dev, bs);
            double ms = 0;
            rc = time_group(b, g, 3, ms);
            if (rc) return rc;
            if (ms < best_ms) { best_ms = ms; cfg = {pn, g->pf_wg_per_cu, bs}; }
          }
        }
      }
      best.push_back(cfg);
    }
  }
  // apply
  b->config.clear();
  for (size_t k = 0; k < b->groups.size(); ++k) {
    auto* g = b->groups[k];
    if (g->pf_nodes != best[k][0]) {
      g->pf_nodes = best[k][0];
      g->pattern_valid = false;
      b->ready = false;
    }
    g->pf_wg_per_cu = best[k][1];
  }
  int rc = batch_ready(b);
  if (rc) return rc;
  for (size_t k = 0; k < b->groups.size(); ++k) {
    auto* g = b->groups[k];
    if (g->path == "plate_fused") {
      auto e = femdev::set_plate_grid(g->dev, best[k][1]);
      if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
      b->config += "group " + std::to_string(k) + ": plate_fused patch_nodes=" + std::to_string(best[k][0]) +
                   " wg_per_cu=" + std::to_string(best[k][1]) + "; ";
    } else if (g->path == "plate_blocks" || g->path == "plate_corners") {
      auto e = femdev::set_plate_block_grid(g->dev, best[k][1]);
      if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
      b->config += "group " + std::to_string(k) + ": " + g->path + " wg_per_cu=" + std::to_string(best[k][1]) + "; ";
    } else {
      if (g->path == "generic") femdev::set_generic_block(g->dev, best[k][2]);
      b->config += "group " + std::to_string(k) + ": " + g->path +
                   (g->path == "generic" ? " block=" + std::to_string(best[k][2]) : std::string()) + "; ";
    }
  }
  if (best_schedule < 0) {
    // the schedule: concurrent branches or one stream, timed on the whole graph
    double best_ms = 1e300;
    for (int sched : {0, 1}) {
      b->schedule = sched;
      rc = batch_capture(b);
      if (rc) return rc;
      double ms = 0;
      rc = time_graph(b, 5, ms);
      if (rc) return rc;
      if (ms < best_ms) { best_ms = ms; best_schedule = sched; }
    }
    cache[key] = {best_schedule, best};
    tune_cache_store(key, best_schedule, best);
  }
  b->schedule = best_schedule;
  b->config += best_schedule == 0 ? "groups concurrent" : "groups serial";
  return batch_capture(b);
}

int fem_batch_assemble(fem_batch* b, void* stream) {
  if (!b) return FEM_ERR_ARGUMENT;
  int rc = batch_ready(b);
  if (rc) return rc;
  if (!b->graph) {
    rc = batch_capture(b);
    if (rc) return rc;
  }
  auto e = femdev::timer_start(b->timer, stream);
  if (e.empty()) e = femdev::graph_launch(b->graph, stream);
  if (e.empty()) e = femdev::timer_end(b->timer, stream);
  if (!e.empty()) return batch_fail(b, FEM_ERR_DEVICE, e);
  b->timed = true;
  return FEM_OK;
}

double fem_batch_last_ms(fem_batch* b) { return b && b->timed ? femdev::last_ms_between(b->timer) : -1.0; }

const char* fem_batch_config(const fem_batch* b) { return b ? b->config.c_str() : ""; }

int fem_batch_mesh_csr(fem_batch* b, size_t mesh, int64_t* n_rows, int64_t* nnz, const int64_t** row_ptr,
                       const int32_t** col_idx, const double** val, int64_t* val_base, int64_t* col_base) {
  if (!b || mesh >= b->meshes.size()) return FEM_ERR_ARGUMENT;
  int rc = batch_ready(b);
  if (rc) return rc;
  const auto& mm = b->meshes[mesh];
  auto* g = b->groups[size_t(mm.group)];
  if (n_rows) *n_rows = int64_t(6 * mm.n_nodes);
  if (nnz) *nnz = mm.nnz;
  if (row_ptr) *row_ptr = femdev::row_ptr(g->dev) + 6 * mm.node_off;
  if (col_idx) *col_idx = femdev::col_idx(g->dev) + mm.val_off;
  if (val) *val = femdev::values(g->dev) + mm.val_off;
  if (val_base) *val_base = mm.val_off;
  if (col_base) *col_base = int64_t(6 * mm.node_off);
  return FEM_OK;
}

int fem_batch_mesh_csr_to_host(fem_batch* b, size_t mesh, int64_t* rp, int32_t* ci, double* v) {
  const int64_t *drp = nullptr;
  const int32_t* dci = nullptr;
  const double* dv = nullptr;
  int64_t nr = 0, nz = 0, vb = 0, cb = 0;
  int rc = fem_batch_mesh_csr(b, mesh, &nr, &nz, &drp, &dci, &dv, &vb, &cb);
  if (rc) return rc;
  std::string e;
  if (rp) {
    e = femdev::copy_to_host(rp, drp, size_t(nr + 1) * sizeof(int64_t));
    for (int64_t i = 0; i <= nr && e.empty(); ++i) rp[i] -= vb;
  }
  if (ci && e.empty()) {
    e = femdev::copy_to_host(ci, dci, size_t(nz) * sizeof(int32_t));
    for (int64_t i = 0; i < nz && e.empty(); ++i) ci[i] -= int32_t(cb);
  }
  if (v && e.empty()) e = femdev::copy_to_host(v, dv, size_t(nz) * sizeof(double));
  return e.empty() ? FEM_OK : batch_fail(b, FEM_ERR_DEVICE, e);
}

}  // extern "C"
