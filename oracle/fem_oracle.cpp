// fem_oracle.cpp -- CPU restatement of the element-stiffness and global-assembly
// path of the Rust crate `finite_element_method` (v0.9.12).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library.  It follows the crate line by line (file:line citations are
// relative to the crate root) and keeps the crate's operation order wherever it
// is cheap to do so: (B^T.C).B per integration point, integration points in the
// crate's order, (T^T.k).T, and sequential `+=` into the global matrix in
// element-insertion order.  It is templated on the float type so that the
// crate's own f32 golden test can be replayed bit-for-bit.
//
// Functions whose arithmetic lives in the absent dependency `extended_matrix`
// (0.9.x, not vendored) are restated from their call-site contracts:
//   * rotation_matrix_to_align_with_vector  -> minimal (Rodrigues) rotation,
//     entries snapped with abs_tol.  Truss and plate stiffness do not depend on
//     the free gauge of this rotation; the beam roll does (parity unpinned for
//     beams not along +x, see the beam section).
//   * SquareMatrix::inverse / determinant (2x2) -> closed-form cofactor formulas.
//   * Vector3::norm -> sqrt((x*x + y*y) + z*z).
//   * add_value -> insert or `+=`; CsrMatrix::from_coo -> rows and columns
//     ascending, duplicates summed in insertion order.
// Those choices only move results at the rounding level (the crate's matrices
// are map-backed, so its summation order is unknown anyway).

#include <algorithm>
#include <array>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr int NODE_DOF = 6;  // src/fem/structs/node.rs:8

// ---------------------------------------------------------------- formatting
// Rust `{:?}` of a float (core::fmt float Debug): shortest round-trip digits,
// exponential form when 0 < |x| < 1e-4 or |x| >= 1e16, else decimal with at
// least one fractional digit.
template <class V>
std::string rust_debug(V x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-inf" : "inf";
  char buf[64];
  auto r = std::to_chars(buf, buf + 64, x, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  bool neg = false;
  if (!s.empty() && s[0] == '-') { neg = true; s = s.substr(1); }
  size_t epos = s.find('e');
  std::string mant = s.substr(0, epos);
  int ex = std::stoi(s.substr(epos + 1));
  std::string digits;
  for (char c : mant) if (c != '.') digits += c;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  V ax = std::fabs(x);
  std::string out = neg || std::signbit(x) ? "-" : "";
  if ((ax != V(0) && ax < V(1e-4)) || ax >= V(1e16)) {
    out += digits.substr(0, 1);
    if (digits.size() > 1) out += "." + digits.substr(1);
    out += "e" + std::to_string(ex);
    return out;
  }
  if (ax == V(0)) return out + "0.0";
  // decimal: value = 0.d1d2... * 10^(ex+1)
  int point = ex + 1;  // digits before the decimal point
  std::string ip, fp;
  if (point <= 0) {
    ip = "0";
    fp = std::string(-point, '0') + digits;
  } else if ((int)digits.size() <= point) {
    ip = digits + std::string(point - digits.size(), '0');
    fp = "0";
  } else {
    ip = digits.substr(0, point);
    fp = digits.substr(point);
  }
  return out + ip + "." + fp;
}

// ---------------------------------------------------------------- helpers
// src/fem/math_functions.rs:3-12
template <class V>
V compare_with_tolerance(V value, V abs_tol) {
  return std::fabs(value) < abs_tol ? V(0) : value;
}
// src/fem/math_functions.rs:14-28: derivative_x(power_func_x, a, x, n)
// = power_func_x(a * n, x, n - 1) = (a * n) * x^(n-1) as a left fold.
template <class V>
V derivative_x(V a, V x, int n) {
  V converted_n = V(0);
  for (int i = 0; i < n; ++i) converted_n += V(1);
  V acc = a * converted_n;
  for (int i = 0; i < n - 1; ++i) acc = acc * x;
  return acc;
}

template <class V>
using Vec3 = std::array<V, 3>;

template <class V>
V norm3(const Vec3<V>& v) { return std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }
template <class V>
Vec3<V> cross3(const Vec3<V>& a, const Vec3<V>& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <class V>
V dot3(const Vec3<V>& a, const Vec3<V>& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// extended_matrix Vector3::rotation_matrix_to_align_with_vector (restated):
// rotation R with R . (a/|a|) = b/|b|, the minimal rotation about a x b.
template <class V>
bool rotation_to_align(const Vec3<V>& a, const Vec3<V>& b, V abs_tol, std::array<V, 9>& R,
                       std::string& err) {
  V na = norm3(a), nb = norm3(b);
  if (na == V(0) || nb == V(0)) {
    err = "Zero length vector could not be aligned!";
    return false;
  }
  Vec3<V> ua{a[0] / na, a[1] / na, a[2] / na};
  Vec3<V> ub{b[0] / nb, b[1] / nb, b[2] / nb};
  Vec3<V> v = cross3(ua, ub);
  V s = norm3(v);
  V c = dot3(ua, ub);
  if (s == V(0)) {
    if (c > V(0)) {
      R = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    } else {
      // half turn about an axis perpendicular to a: R = 2 n n^T - I
      Vec3<V> e{V(1), V(0), V(0)};
      if (std::fabs(ua[0]) > std::fabs(ua[1]) || std::fabs(ua[0]) > std::fabs(ua[2])) {
        e = std::fabs(ua[1]) <= std::fabs(ua[2]) ? Vec3<V>{V(0), V(1), V(0)}
                                                 : Vec3<V>{V(0), V(0), V(1)};
      }
      Vec3<V> n = cross3(ua, e);
      V nn = norm3(n);
      n = {n[0] / nn, n[1] / nn, n[2] / nn};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = V(2) * n[i] * n[j] - (i == j ? V(1) : V(0));
    }
  } else {
    Vec3<V> k{v[0] / s, v[1] / s, v[2] / s};
    // R = c I + s [k]x + (1 - c) k k^T
    V K[9] = {V(0), -k[2], k[1], k[2], V(0), -k[0], -k[1], k[0], V(0)};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        R[i * 3 + j] = (i == j ? c : V(0)) + s * K[i * 3 + j] + (V(1) - c) * k[i] * k[j];
  }
  for (auto& r : R) r = compare_with_tolerance(r, abs_tol);
  return true;
}

// dense row-major helpers (ascending-k summation)
template <class V>
std::vector<V> matmul(const std::vector<V>& A, const std::vector<V>& B, int n, int m, int p) {
  std::vector<V> C(size_t(n) * p, V(0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < p; ++j) {
      V acc = V(0);
      for (int k = 0; k < m; ++k) acc += A[size_t(i) * m + k] * B[size_t(k) * p + j];
      C[size_t(i) * p + j] = acc;
    }
  return C;
}
template <class V>
std::vector<V> transpose(const std::vector<V>& A, int n, int m) {
  std::vector<V> T(size_t(n) * m);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) T[size_t(j) * n + i] = A[size_t(i) * m + j];
  return T;
}
// (T^T . k) . T with T = block-diag of `nb` copies of R (src/fem/structs/*.rs
// compose_rotation_matrix; methods_for_*_data_handle.rs transformed_local_...)
template <class V>
std::vector<V> rotate_tkt(const std::vector<V>& k, const std::array<V, 9>& R, int n) {
  std::vector<V> T(size_t(n) * n, V(0));
  for (int b = 0; b < n / 3; ++b)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[size_t(3 * b + i) * n + 3 * b + j] = R[i * 3 + j];
  auto Tt = transpose(T, n, n);
  return matmul(matmul(Tt, k, n, n, n), T, n, n, n);
}

// structural non-zero pattern of (T^T k T) from the zero pattern of R and the
// structural pattern of the local k
template <class V>
std::vector<uint8_t> rotate_pattern(const std::vector<uint8_t>& P, const std::array<V, 9>& R, int n) {
  std::vector<uint8_t> T(size_t(n) * n, 0);
  for (int b = 0; b < n / 3; ++b)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[size_t(3 * b + i) * n + 3 * b + j] = R[i * 3 + j] != V(0);
  std::vector<uint8_t> out(size_t(n) * n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      bool nz = false;
      for (int k = 0; k < n && !nz; ++k) {
        if (!T[size_t(k) * n + i]) continue;
        for (int l = 0; l < n; ++l)
          if (P[size_t(k) * n + l] && T[size_t(l) * n + j]) { nz = true; break; }
      }
      out[size_t(i) * n + j] = nz;
    }
  return out;
}

// ---------------------------------------------------------------- model
template <class V>
struct Node {
  size_t index;
  V x, y, z;
};

template <class V>
struct Truss {
  uint32_t n1, n2;
  V E, A;
  bool has_a2;
  V A2;
  std::array<V, 9> R;
};
template <class V>
struct Beam {
  uint32_t n1, n2;
  V E, nu, A, i11p, i22p, it, k;
  std::array<V, 9> R;
};
template <class V>
struct Plate {
  uint32_t n[4];
  V E, nu, t, k;
  std::array<V, 9> R;
  std::array<V, 4> ip_rs;  // g for (r, s) signs, see plate.rs:1066-1091
};

struct Triplet {
  int64_t row, col;
  size_t seq;
};

template <class V>
struct Model {
  V rel_tol, abs_tol;
  uint32_t nodes_number;
  std::unordered_map<uint32_t, Node<V>> nodes;
  std::unordered_map<uint64_t, std::vector<uint32_t>> coord_index;  // exact-coordinate lookup
  size_t nodes_count = 0;
  std::unordered_map<uint32_t, Truss<V>> trusses;
  std::unordered_map<uint32_t, Beam<V>> beams;
  std::unordered_map<uint32_t, Plate<V>> plates;
  std::unordered_map<uint64_t, uint32_t> truss_pairs, beam_pairs;      // sorted node pair -> element
  std::unordered_map<uint32_t, std::vector<uint32_t>> node_plates;     // node -> plates
  // global matrix contributions in add_value order (src/fem/methods_for_*_data_handle.rs)
  std::vector<int64_t> t_row, t_col;
  std::vector<V> t_val;
  std::vector<uint8_t> t_structural;  // 1 if the entry is in the element's structural pattern
  std::vector<uint8_t> t_nonzero;     // 1 if the crate would call add_value (value != 0)
  std::vector<uint16_t> t_local;      // position li * n + lj of the entry in its element matrix
  struct ERange {
    int type;
    uint32_t number;
    size_t t0, t1;
  };
  std::vector<ERange> erange;  // elements in insertion order and their triplet ranges
  std::string last_error;
  // boundary conditions (methods_for_bc_data_handle.rs): per global dof
  std::vector<uint8_t> constrained;
  std::vector<V> displacement, force;
  // sparse separation (methods_for_separate_stiffness_matrix.rs:217-320)
  std::vector<int64_t> aa_idx, bb_idx;
  std::vector<int64_t> sep_rp[4];
  std::vector<int32_t> sep_ci[4];
  std::vector<double> sep_v[4];
  // exported CSR
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col_idx;
  std::vector<double> val;
};

template <class V>
uint64_t coord_key(V x, V y, V z) {
  auto h = [](V v) {
    double d = double(v);
    if (d == 0.0) d = 0.0;  // -0 == 0
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return u;
  };
  return h(x) * 0x9E3779B97F4A7C15ull ^ (h(y) + 0x7F4A7C15ull) * 0xC2B2AE3D27D4EB4Full ^ h(z);
}

// src/fem/methods_for_node_data_handle.rs:39-79
template <class V>
bool add_node(Model<V>& m, uint32_t number, V x, V y, V z) {
  size_t node_index = m.nodes_count;
  if (m.nodes_number == 0 || node_index > size_t(m.nodes_number) - 1) {
    m.last_error = "Nodes number could not be greater than " + std::to_string(m.nodes_number) + "!";
    return false;
  }
  if (m.nodes.count(number)) {
    m.last_error = "Node with number " + std::to_string(number) + " already exists!";
    return false;
  }
  auto key = coord_key(x, y, z);
  auto it = m.coord_index.find(key);
  if (it != m.coord_index.end()) {
    for (uint32_t other : it->second) {
      const auto& nd = m.nodes.at(other);
      if (nd.x == x && nd.y == y && nd.z == z) {
        m.last_error = "Node with coordinates x: " + rust_debug(x) + ", y: " + rust_debug(y) +
                       ", z: " + rust_debug(z) + " already exists!";
        return false;
      }
    }
  }
  m.nodes[number] = Node<V>{node_index, x, y, z};
  m.coord_index[key].push_back(number);
  m.nodes_count += 1;
  return true;
}

template <class V>
bool check_node_exist(Model<V>& m, uint32_t number) {  // methods_for_node_data_handle.rs:81-86
  if (!m.nodes.count(number)) {
    m.last_error = "Node with number " + std::to_string(number) + " does not exist!";
    return false;
  }
  return true;
}

template <class V>
Vec3<V> coords(const Model<V>& m, uint32_t n) {
  const auto& nd = m.nodes.at(n);
  return {nd.x, nd.y, nd.z};
}
// bar_2n_element_functions.rs:8-33
template <class V>
Vec3<V> element_vector(const Model<V>& m, uint32_t n1, uint32_t n2) {
  auto a = coords(m, n1), b = coords(m, n2);
  return {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
}

// ---------------------------------------------------------------- bar kinematics
// bar_2n_element_functions.rs:35-113
template <class V>
V bar_jacobian(V length, V r) {
  V x_1 = V(-1) * length / V(2);
  V x_2 = length / V(2);
  return derivative_x(x_1 * V(0.5f), V(0), 0) - derivative_x(x_1 * V(0.5f), r, 1) +
         derivative_x(x_2 * V(0.5f), V(0), 0) + derivative_x(x_2 * V(0.5f), r, 1);
}
template <class V> V bar_h1(V r) { return V(0.5f) * (V(1) - r); }
template <class V> V bar_h2(V r) { return V(0.5f) * (V(1) + r); }
template <class V> V bar_dh1(V r) { return derivative_x(V(0.5f), V(0), 0) - derivative_x(V(0.5f), r, 1); }
template <class V> V bar_dh2(V r) { return derivative_x(V(0.5f), V(0), 0) + derivative_x(V(0.5f), r, 1); }

// ---------------------------------------------------------------- truss
// src/fem/structs/truss.rs:95-189
template <class V>
std::vector<V> truss_local_k(const Truss<V>& e, V length) {
  const int n = 6;
  V r = V(0), alpha = V(2);  // truss.rs:244
  V inv_j = V(1) / bar_jacobian(length, r);
  V det_j = bar_jacobian(length, r);
  std::vector<V> B = {bar_dh1(r), V(0), V(0), bar_dh2(r), V(0), V(0)};
  for (auto& b : B) b = b * inv_j;
  V area = e.has_a2 ? (e.A2 - e.A) / V(2) * r + e.A - (e.A2 - e.A) / V(2) * V(-1) : e.A;
  V c = area * e.E;
  V scale = c * det_j * alpha;
  std::vector<V> k(n * n, V(0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) k[i * n + j] = V(0) + (V(0) + B[i] * B[j]) * scale;
  return k;
}
std::vector<uint8_t> truss_local_pattern() {
  std::vector<uint8_t> P(36, 0);
  for (int i : {0, 3})
    for (int j : {0, 3}) P[i * 6 + j] = 1;
  return P;
}

// ---------------------------------------------------------------- beam
// src/fem/structs/beam.rs:119-160
template <class V>
void principal_moments(V i11, V i22, V i12, V rel_tol, V& i11_p, V& i22_p, V& angle) {
  const V PI_F32 = V(float(3.14159265358979323846));  // std::f32::consts::PI
  if (i11 != i22) {
    angle = std::atan(V(2) * i12 / (i22 - i11)) / V(2);
  } else {
    V i11_mod, i22_mod;
    if (i22 < i11) {
      i11_mod = i11;
      i22_mod = (std::fabs(i22) - std::fabs(i22) * rel_tol) * i22 / std::fabs(i22);
    } else {
      i11_mod = (std::fabs(i11) - std::fabs(i11) * rel_tol) * i11 / std::fabs(i11);
      i22_mod = i22;
    }
    angle = std::atan(V(2) * i12 / (i22_mod - i11_mod)) / V(2);
  }
  auto sq = [](V v) { return v * v; };
  i11_p = i11 * sq(std::cos(angle)) + i22 * sq(std::sin(angle)) - i12 * std::sin(V(2) * angle);
  i22_p = i11 * sq(std::sin(angle)) + i22 * sq(std::cos(angle)) + i12 * std::sin(V(2) * angle);
  int i = 1;
  while (i11_p < i22_p) {
    angle = (std::atan(V(2) * i12 / (i22 - i11)) + PI_F32 * V(float(i))) / V(2);
    i11_p = i11 * sq(std::cos(angle)) + i22 * sq(std::sin(angle)) - i12 * std::sin(V(2) * angle);
    i22_p = i11 * sq(std::sin(angle)) + i22 * sq(std::cos(angle)) + i12 * std::sin(V(2) * angle);
    i += 1;
    if (i > 64) break;  // NaN guard; the crate's loop cannot run more than twice for finite input
  }
}

// src/fem/structs/beam.rs:162-258
template <class V>
bool beam_rotation(const Vec3<V>& d, const Vec3<V>& v1, V angle, V rel_tol, V abs_tol,
                   std::array<V, 9>& out, std::string& err) {
  (void)rel_tol;
  V length = norm3(d);
  std::array<V, 9> R0;
  if (!rotation_to_align(d, Vec3<V>{length, V(0), V(0)}, abs_tol, R0, err)) return false;
  // projection_perpendicular_to_vector: v1 - (v1.d / d.d) d
  V f = dot3(v1, d) / dot3(d, d);
  Vec3<V> p{v1[0] - f * d[0], v1[1] - f * d[1], v1[2] - f * d[2]};
  Vec3<V> tp;
  for (int i = 0; i < 3; ++i) tp[i] = (R0[i * 3 + 0] * p[0] + R0[i * 3 + 1] * p[1]) + R0[i * 3 + 2] * p[2];
  // cosine_angle_between_vectors(z, tp)
  Vec3<V> zz{V(0), V(0), V(1)};
  V cosang = dot3(zz, tp) / (norm3(zz) * norm3(tp));
  V roll = std::acos(cosang);
  V total_angle = angle + roll;
  V x = d[0], y = d[1], z = d[2];
  V c_x = compare_with_tolerance(x / length, abs_tol);
  V c_y = compare_with_tolerance(y / length, abs_tol);
  V c_z = compare_with_tolerance(z / length, abs_tol);
  V c_xz = compare_with_tolerance(std::sqrt(c_x * c_x + c_z * c_z), abs_tol);
  V c = compare_with_tolerance(std::cos(total_angle), abs_tol);
  V s = compare_with_tolerance(std::sin(total_angle), abs_tol);
  bool nz = c_xz != V(0);
  V r11 = nz ? c_x : V(0);
  V r12 = c_y;
  V r13 = nz ? c_z : V(0);
  V r21 = nz ? (V(-1) * c_x * c_y * c - c_z * s) / c_xz : V(-1) * c_y * c;
  V r22 = nz ? c_xz * c : V(0);
  V r23 = nz ? (V(-1) * c_y * c_z * c + c_x * s) / c_xz : s;
  V r31 = nz ? (c_x * c_y * s - c_z * c) / c_xz : c_y * s;
  V r32 = nz ? V(-1) * c_xz * s : V(0);
  V r33 = nz ? (c_y * c_z * s + c_x * c) / c_xz : c;
  out = {r11, r12, r13, r21, r22, r23, r31, r32, r33};
  return true;
}

// src/fem/structs/beam.rs:260-605 (one integration point r = 0, alpha = 2)
// parts (test hook): bit t keeps the rank-one term t (u, v, w, tu, tv, tw)
template <class V>
std::vector<V> beam_local_k(const Beam<V>& e, V length, unsigned parts = 63) {
  const int n = 12;
  V r = V(0), alpha = V(2);
  V inv_j = V(1) / bar_jacobian(length, r);
  V det_j = bar_jacobian(length, r);
  auto row = [&](std::initializer_list<std::pair<int, V>> lhs, std::initializer_list<std::pair<int, V>> rhs,
                 bool scale_lhs) {
    std::vector<V> b(n, V(0));
    for (auto& p : lhs) b[p.first] = scale_lhs ? p.second * inv_j : p.second;
    std::vector<V> out(n, V(0));
    for (int i = 0; i < n; ++i) out[i] = b[i];
    if (rhs.size()) {
      std::vector<V> h(n, V(0));
      for (auto& p : rhs) h[p.first] = p.second;
      for (int i = 0; i < n; ++i) out[i] = b[i] - h[i];  // Matrix::subtract
    }
    return out;
  };
  std::vector<V> bu = row({{0, bar_dh1(r)}, {6, bar_dh2(r)}}, {}, true);
  std::vector<V> bv = row({{1, bar_dh1(r)}, {7, bar_dh2(r)}}, {{5, bar_h1(r)}, {11, bar_h2(r)}}, true);
  std::vector<V> bw = row({{2, bar_dh1(r)}, {8, bar_dh2(r)}}, {{4, bar_h1(r)}, {10, bar_h2(r)}}, true);
  std::vector<V> btu = row({{3, bar_dh1(r)}, {9, bar_dh2(r)}}, {}, true);
  std::vector<V> btv = row({{4, bar_dh1(r)}, {10, bar_dh2(r)}}, {}, true);
  std::vector<V> btw = row({{5, bar_dh1(r)}, {11, bar_dh2(r)}}, {}, true);
  V G = e.E / (V(2) * (V(1) + e.nu));
  V cs[6] = {e.A * e.E, G * e.A * e.k, G * e.A * e.k, G * e.it, e.E * e.i22p, e.E * e.i11p};
  for (int t = 0; t < 6; ++t)
    if (!((parts >> t) & 1u)) cs[t] = V(0);
  const std::vector<V>* Bs[6] = {&bu, &bv, &bw, &btu, &btv, &btw};
  std::vector<V> kip(n * n, V(0));
  for (int t = 0; t < 6; ++t) {
    V scale = cs[t] * det_j * alpha;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        V term = (V(0) + (*Bs[t])[i] * (*Bs[t])[j]) * scale;
        kip[i * n + j] = t == 0 ? term : kip[i * n + j] + term;
      }
  }
  std::vector<V> k(n * n);
  for (int i = 0; i < n * n; ++i) k[i] = V(0) + kip[i];
  return k;
}
std::vector<uint8_t> beam_local_pattern() {
  std::vector<uint8_t> P(144, 0);
  auto set_group = [&](std::initializer_list<int> dofs) {
    for (int a : dofs)
      for (int b : dofs) P[a * 12 + b] = 1;
  };
  set_group({0, 6});            // u
  set_group({1, 5, 7, 11});     // v, theta_w
  set_group({2, 4, 8, 10});     // w, theta_v
  set_group({3, 9});            // theta_u
  return P;
}

// Structural pattern of the beam's global K_e.  Generic rule (T^T P T) except
// for the translational 3x3 blocks: the v and w shear terms carry bitwise equal
// coefficients, so their sum is kv (r1 r1^T + r2 r2^T) = kv (I - r0 r0^T), whose
// off-diagonal entries vanish identically when r0_i r0_j = 0 (r0 = beam axis).
template <class V>
std::vector<uint8_t> beam_global_pattern(const std::array<V, 9>& R) {
  auto P = rotate_pattern(beam_local_pattern(), R, 12);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          if (i != j) P[(6 * a + i) * 12 + 6 * b + j] = R[i] != V(0) && R[j] != V(0);
  return P;
}

// ---------------------------------------------------------------- quadrilateral
// src/fem/quadrilateral_4n_element_functions.rs:252-338
template <class V>
V quad_dx_dr(V x1, V x2, V x3, V x4, V r, V s) {
  V q = V(0.25f);
  return derivative_x(x1 * q, V(0), 0) + derivative_x(x1 * q * s, V(0), 0) + derivative_x(x1 * q, r, 1) +
         derivative_x(x1 * q * s, r, 1) + derivative_x(x2 * q, V(0), 0) + derivative_x(x2 * q * s, V(0), 0) -
         derivative_x(x2 * q, r, 1) - derivative_x(x2 * q * s, r, 1) + derivative_x(x3 * q, V(0), 0) -
         derivative_x(x3 * q * s, V(0), 0) - derivative_x(x3 * q, r, 1) + derivative_x(x3 * q * s, r, 1) +
         derivative_x(x4 * q, V(0), 0) - derivative_x(x4 * q * s, V(0), 0) + derivative_x(x4 * q, r, 1) -
         derivative_x(x4 * q * s, r, 1);
}
template <class V>
V quad_dx_ds(V x1, V x2, V x3, V x4, V r, V s) {
  V q = V(0.25f);
  return derivative_x(x1 * q, V(0), 0) + derivative_x(x1 * q, s, 1) + derivative_x(x1 * q * r, V(0), 0) +
         derivative_x(x1 * q * r, s, 1) + derivative_x(x2 * q, V(0), 0) + derivative_x(x2 * q, s, 1) -
         derivative_x(x2 * q * r, V(0), 0) - derivative_x(x2 * q * r, s, 1) + derivative_x(x3 * q, V(0), 0) -
         derivative_x(x3 * q, s, 1) - derivative_x(x3 * q * r, V(0), 0) + derivative_x(x3 * q * r, s, 1) +
         derivative_x(x4 * q, V(0), 0) - derivative_x(x4 * q, s, 1) + derivative_x(x4 * q * r, V(0), 0) -
         derivative_x(x4 * q * r, s, 1);
}
// quad:505-583 shape-function derivatives
template <class V>
void quad_dh(V r, V s, V dr[4], V ds[4]) {
  V q = V(0.25f);
  dr[0] = derivative_x(q, V(0), 0) + derivative_x(q * s, V(0), 0) + derivative_x(q, r, 1) + derivative_x(q * s, r, 1);
  dr[1] = derivative_x(q, V(0), 0) + derivative_x(q * s, V(0), 0) - derivative_x(q, r, 1) - derivative_x(q * s, r, 1);
  dr[2] = derivative_x(q, V(0), 0) - derivative_x(q * s, V(0), 0) - derivative_x(q, r, 1) + derivative_x(q * s, r, 1);
  dr[3] = derivative_x(q, V(0), 0) - derivative_x(q * s, V(0), 0) + derivative_x(q, r, 1) - derivative_x(q * s, r, 1);
  ds[0] = derivative_x(q, V(0), 0) + derivative_x(q, s, 1) + derivative_x(q * r, V(0), 0) + derivative_x(q * r, s, 1);
  ds[1] = derivative_x(q, V(0), 0) + derivative_x(q, s, 1) - derivative_x(q * r, V(0), 0) - derivative_x(q * r, s, 1);
  ds[2] = derivative_x(q, V(0), 0) - derivative_x(q, s, 1) - derivative_x(q * r, V(0), 0) + derivative_x(q * r, s, 1);
  ds[3] = derivative_x(q, V(0), 0) - derivative_x(q, s, 1) + derivative_x(q * r, V(0), 0) - derivative_x(q * r, s, 1);
}

// quad:131-171 frame of a quadrilateral from nodes 2, 3, 4
template <class V>
bool quad_rotation(const Vec3<V>& p2, const Vec3<V>& p3, const Vec3<V>& p4, V abs_tol, std::array<V, 9>& R,
                   std::string& err) {
  Vec3<V> e34{p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
  Vec3<V> e32{p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
  Vec3<V> nrm = cross3(e34, e32);
  V len = norm3(nrm);
  return rotation_to_align(nrm, Vec3<V>{V(0), V(0), len}, abs_tol, R, err);
}

// quad:14-86
template <class V>
bool quad_on_line(const Vec3<V>& p1, const Vec3<V>& p2, const Vec3<V>& p3, const Vec3<V>& p4) {
  auto sub = [](const Vec3<V>& a, const Vec3<V>& b) { return Vec3<V>{a[0] - b[0], a[1] - b[1], a[2] - b[2]}; };
  std::array<std::array<Vec3<V>, 2>, 4> pairs = {{{sub(p2, p1), sub(p4, p1)},
                                                  {sub(p1, p2), sub(p3, p2)},
                                                  {sub(p2, p3), sub(p4, p3)},
                                                  {sub(p3, p4), sub(p1, p4)}}};
  for (auto& pr : pairs)
    if (norm3(cross3(pr[0], pr[1])) == V(0)) return true;
  return false;
}
// quad:88-129
template <class V>
bool quad_on_plane(const Vec3<V>& p1, const Vec3<V>& p2, const Vec3<V>& p3, const Vec3<V>& p4, V abs_tol) {
  Vec3<V> v32{p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
  Vec3<V> v34{p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
  Vec3<V> n = cross3(v32, v34);
  V a = n[0], b = n[1], c = n[2];
  V d = V(-1) * (a * p3[0] + b * p3[1] + c * p3[2]);
  return compare_with_tolerance(a * p1[0] + b * p1[1] + c * p1[2] + d, abs_tol) == V(0);
}

// src/fem/convex_hull_on_plane.rs:14-214
template <class V>
struct HullPoint {
  uint32_t n;
  V x, y;
};
template <class V>
V hull_angle(const HullPoint<V>& p) {
  // cosine between (0,0)->(1,0) and (0,0)->(x,y), then acos in degrees
  V lhs = std::sqrt((V(0) - V(1)) * (V(0) - V(1)) + (V(0) - V(0)) * (V(0) - V(0)));
  V rhs = std::sqrt((V(0) - p.x) * (V(0) - p.x) + (V(0) - p.y) * (V(0) - p.y));
  V sp = (V(0) - V(1)) * (V(0) - p.x) + (V(0) - V(0)) * (V(0) - p.y);
  V deg = std::acos(sp / (lhs * rhs)) * (V(180) / V(3.14159265358979323846));
  return deg;
}
template <class V>
bool hull_lt(const HullPoint<V>& a, const HullPoint<V>& b) { return hull_angle(a) < hull_angle(b); }
template <class V>
bool hull_gt(const HullPoint<V>& a, const HullPoint<V>& b) { return hull_angle(a) > hull_angle(b); }
template <class V>
long hull_partition(std::vector<HullPoint<V>>& arr, size_t off, long low, long high) {
  long pivot = high;
  long store = low - 1;
  long last = high;
  while (true) {
    store += 1;
    while (hull_lt(arr[off + store], arr[off + pivot])) store += 1;
    last -= 1;
    while (last >= 0 && hull_gt(arr[off + last], arr[off + pivot])) last -= 1;
    if (store >= last) break;
    std::swap(arr[off + store], arr[off + last]);
  }
  std::swap(arr[off + store], arr[off + pivot]);
  return store;
}
template <class V>
void hull_qsort(std::vector<HullPoint<V>>& arr, size_t off, long low, long high) {
  if (low < high) {
    long p = hull_partition(arr, off, low, high);
    hull_qsort(arr, off, low, p - 1);
    hull_qsort(arr, off, p + 1, high);
  }
}
template <class V>
std::vector<HullPoint<V>> convex_hull_on_plane(std::vector<HullPoint<V>> u) {
  V shift_x = u[0].x, min_y = u[0].y;
  size_t min_pos = 0;
  for (size_t i = 0; i < u.size(); ++i)
    if (u[i].y < min_y) { shift_x = u[i].x; min_y = u[i].y; min_pos = i; }
  std::swap(u[0], u[min_pos]);
  for (auto& p : u) { p.x -= shift_x; p.y -= min_y; }
  hull_qsort(u, 1, 0, long(u.size() - 1) - 1);
  size_t i = 0;
  while (i + 2 < u.size()) {
    const auto &a = u[i], &b = u[i + 1], &c = u[i + 2];
    V area = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
    if (area <= V(0)) u.erase(u.begin() + i + 1);
    else i += 1;
  }
  for (auto& p : u) { p.x += shift_x; p.y += min_y; }
  return u;
}
// quad:173-250
template <class V>
bool quad_hull(const uint32_t num[4], const Vec3<V> pts[4], V abs_tol, std::vector<HullPoint<V>>& hull_out,
               std::string& err) {
  std::array<V, 9> R;
  if (!quad_rotation(pts[1], pts[2], pts[3], abs_tol, R, err)) return false;
  auto tr = [&](const Vec3<V>& p) {
    Vec3<V> d{p[0] - pts[2][0], p[1] - pts[2][1], p[2] - pts[2][2]};
    Vec3<V> o;
    for (int i = 0; i < 3; ++i) o[i] = (R[i * 3] * d[0] + R[i * 3 + 1] * d[1]) + R[i * 3 + 2] * d[2];
    return o;
  };
  auto t1 = tr(pts[0]), t2 = tr(pts[1]), t4 = tr(pts[3]);
  std::vector<HullPoint<V>> hp = {{num[0], t1[0], t1[1]}, {num[1], t2[0], t2[1]}, {num[2], V(0), V(0)},
                                  {num[3], t4[0], t4[1]}};
  hull_out = convex_hull_on_plane(hp);
  return true;
}
template <class V>
bool quad_convex(const uint32_t num[4], const Vec3<V> pts[4], V abs_tol, size_t& hull_len, std::string& err) {
  std::vector<HullPoint<V>> h;
  if (!quad_hull(num, pts, abs_tol, h, err)) return false;
  hull_len = h.size();
  return true;
}

// local 2D coordinates of nodes 1, 2, 4 relative to node 3 (quad:340-378)
template <class V>
struct QuadLocal {
  V x1, y1, x2, y2, x4, y4;
};
template <class V>
QuadLocal<V> quad_local(const Model<V>& m, const Plate<V>& e) {
  auto p3 = coords(m, e.n[2]);
  auto tr = [&](uint32_t nn) {
    auto p = coords(m, nn);
    Vec3<V> d{p[0] - p3[0], p[1] - p3[1], p[2] - p3[2]};
    Vec3<V> o;
    for (int i = 0; i < 3; ++i) o[i] = (e.R[i * 3] * d[0] + e.R[i * 3 + 1] * d[1]) + e.R[i * 3 + 2] * d[2];
    return o;
  };
  auto a = tr(e.n[0]), b = tr(e.n[1]), c = tr(e.n[3]);
  return {a[0], a[1], b[0], b[1], c[0], c[1]};
}

// quad:380-503 Jacobian, inverse and determinant at (r, s)
template <class V>
void quad_jacobian(const QuadLocal<V>& q, V r, V s, V J[4]) {
  J[0] = quad_dx_dr(q.x1, q.x2, V(0), q.x4, r, s);
  J[1] = quad_dx_dr(q.y1, q.y2, V(0), q.y4, r, s);  // dy_dr has the dx_dr form
  J[2] = quad_dx_ds(q.x1, q.x2, V(0), q.x4, r, s);
  J[3] = quad_dx_ds(q.y1, q.y2, V(0), q.y4, r, s);
}

// ---------------------------------------------------------------- plate
// src/fem/structs/plate.rs:160-724
// parts (test hook for reference gradients): bit 0 membrane, 1 bending,
// 2 shear, 3 drilling; the default sums all four exactly as the crate does
// B_mem, B_bend (3 x 24) and the MITC4 B_shear (2 x 24) of the local frame at
// (r, s), and det J (plate.rs:160-511, quad:252-653)
template <class V>
void plate_b(const QuadLocal<V>& q, V r, V s, std::vector<V>& bm, std::vector<V>& bb, std::vector<V>& bs, V& det) {
  const int n = 24;
  V J[4];
  quad_jacobian(q, r, s, J);
  det = J[0] * J[3] - J[1] * J[2];
  V inv[4] = {J[3] / det, V(-1) * J[1] / det, V(-1) * J[2] / det, J[0] / det};
  V dr[4], ds[4];
  quad_dh(r, s, dr, ds);
  V dx[4], dy[4];
  for (int a = 0; a < 4; ++a) {
    dx[a] = inv[0] * dr[a] + inv[1] * ds[a];
    dy[a] = inv[2] * dr[a] + inv[3] * ds[a];
  }
  // B_mem, B_bend (3 x 24)
  bm.assign(3 * n, V(0));
  bb.assign(3 * n, V(0));
  bs.assign(2 * n, V(0));
  for (int a = 0; a < 4; ++a) {
    bm[0 * n + 6 * a + 0] = dx[a];
    bm[1 * n + 6 * a + 1] = dy[a];
    bm[2 * n + 6 * a + 0] = dy[a];
    bm[2 * n + 6 * a + 1] = dx[a];
    bb[0 * n + 6 * a + 4] = V(-1) * dx[a];
    bb[1 * n + 6 * a + 3] = dy[a];
    bb[2 * n + 6 * a + 3] = dx[a];
    bb[2 * n + 6 * a + 4] = V(-1) * dy[a];
  }
  // B_shear (plate.rs:392-511), MITC4
  V x1 = q.x1, y1 = q.y1, x2 = q.x2, y2 = q.y2, x3 = V(0), y3 = V(0), x4 = q.x4, y4 = q.y4;
  V a_x = x1 - x2 - x3 + x4, b_x = x1 - x2 + x3 - x4, c_x = x1 + x2 - x3 - x4;
  V a_y = y1 - y2 - y3 + y4, b_y = y1 - y2 + y3 - y4, c_y = y1 + y2 - y3 - y4;
  auto sq = [](V v) { return v * v; };
  V mr = std::sqrt(sq(c_x + r * b_x) + sq(c_y + r * b_y)) / (V(8) * det);
  V ms = std::sqrt(sq(a_x + s * b_x) + sq(a_y + s * b_y)) / (V(8) * det);
  V one = V(1);
  V* R0 = &bs[0];
  V* R1 = &bs[n];
  R0[2] = (one + s) / V(2) * mr;
  R0[3] = (one + s) * V(-1) * (y1 - y2) / V(4) * mr;
  R0[4] = (one + s) * (x1 - x2) / V(4) * mr;
  R0[8] = V(-1) * (one + s) / V(2) * mr;
  R0[9] = (one + s) * V(-1) * (y1 - y2) / V(4) * mr;
  R0[10] = (one + s) * (x1 - x2) / V(4) * mr;
  R0[14] = V(-1) * (one - s) / V(2) * mr;
  R0[15] = (one - s) * V(-1) * (y4 - y3) / V(4) * mr;
  R0[16] = (one - s) * (x4 - x3) / V(4) * mr;
  R0[20] = (one - s) / V(2) * mr;
  R0[21] = (one - s) * V(-1) * (y4 - y3) / V(4) * mr;
  R0[22] = (one - s) * (x4 - x3) / V(4) * mr;
  R1[2] = (one + r) / V(2) * ms;
  R1[3] = (one + r) * V(-1) * (y1 - y4) / V(4) * ms;
  R1[4] = (one + r) * (x1 - x4) / V(4) * ms;
  R1[8] = (one - r) / V(2) * ms;
  R1[9] = (one - r) * V(-1) * (y2 - y3) / V(4) * ms;
  R1[10] = (one - r) * (x2 - x3) / V(4) * ms;
  R1[14] = V(-1) * (one - r) / V(2) * ms;
  R1[15] = (one - r) * V(-1) * (y2 - y3) / V(4) * ms;
  R1[16] = (one - r) * (x2 - x3) / V(4) * ms;
  R1[20] = V(-1) * (one + r) / V(2) * ms;
  R1[21] = (one + r) * V(-1) * (y1 - y4) / V(4) * ms;
  R1[22] = (one + r) * (x1 - x4) / V(4) * ms;
}

template <class V>
std::vector<V> plate_local_k(const Model<V>& m, const Plate<V>& e, unsigned parts = 15) {
  const int n = 24;
  QuadLocal<V> q = quad_local(m, e);
  V ips[4][2] = {{e.ip_rs[0], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[1]}, {e.ip_rs[0], e.ip_rs[1]}};
  std::vector<V> k(n * n, V(0));
  for (int ip = 0; ip < 4; ++ip) {
    V r = ips[ip][0], s = ips[ip][1];
    V alpha = V(1) * V(1);
    V det;
    std::vector<V> bm, bb, bs;
    plate_b(q, r, s, bm, bb, bs, det);
    // constitutive matrices (plate.rs:532-636)
    V nu = e.nu;
    V cm_mul = e.E * e.t / (V(1) - nu * nu);
    V cb_mul = e.E * (e.t * e.t * e.t) / (V(12) * (V(1) - nu * nu));
    V cs_mul = e.E * e.t * e.k / (V(2) * (V(1) + nu));
    std::vector<V> base = {V(1), nu, V(0), nu, V(1), V(0), V(0), V(0), (V(1) - nu) / V(2)};
    std::vector<V> cm(9), cb(9);
    for (int i = 0; i < 9; ++i) { cm[i] = base[i] * cm_mul; cb[i] = base[i] * cb_mul; }
    std::vector<V> cs = {V(1) * (cs_mul * alpha), V(0) * (cs_mul * alpha), V(0) * (cs_mul * alpha),
                         V(1) * (cs_mul * alpha)};
    auto term = [&](const std::vector<V>& B, const std::vector<V>& C, int rows) {
      auto Bt = transpose(B, rows, n);
      return matmul(matmul(Bt, C, n, rows, rows), B, n, rows, n);
    };
    auto km = term(bm, cm, 3), kb = term(bb, cb, 3), ks = term(bs, cs, 2);
    V w = det * alpha;
    for (int i = 0; i < n * n; ++i) {
      V kip = parts == 15 ? (km[i] * w + kb[i] * w) + ks[i] * w
                          : ((parts & 1 ? km[i] * w : V(0)) + (parts & 2 ? kb[i] * w : V(0))) +
                                (parts & 4 ? ks[i] * w : V(0));
      k[i] = k[i] + kip;
    }
  }
  if (parts & 8)
    for (int i = 0; i < 4; ++i) k[(6 * i + 5) * n + 6 * i + 5] += V(1.0f);  // KROT6, plate.rs:716-721
  return k;
}
// ---------------------------------------------------------------- magnitudes
// Term magnitudes of the element matrices (test infrastructure for the parity
// tolerance): every product of the stiffness formulas above taken with the
// absolute value of each factor and summed, i.e. sum_terms prod |factor|, and
// rotated as |T|^T |k| |T|.  Any two evaluations of an entry that sum the same
// terms in different orders differ by at most 2 gamma_n times this magnitude,
// gamma_n ~ n u for chains of n roundings (Higham, Accuracy and Stability of
// Numerical Algorithms, 2nd ed., sec. 3.1), whatever cancellation the signed
// sum has.  Geometric factors (local coordinates, Jacobian, frame) enter as
// computed.
template <class V>
std::vector<V> rotate_tkt_abs(const std::vector<V>& k, const std::array<V, 9>& R, int n) {
  std::array<V, 9> Ra;
  for (int i = 0; i < 9; ++i) Ra[i] = std::fabs(R[i]);
  return rotate_tkt(k, Ra, n);
}
template <class V>
std::vector<V> abs_all(std::vector<V> a) {
  for (auto& x : a) x = std::fabs(x);
  return a;
}
// beam_local_k with every term |B_t[i]| |B_t[j]| |scale_t|
template <class V>
std::vector<V> beam_local_k_abs(const Beam<V>& e, V length) {
  const int n = 12;
  V r = V(0), alpha = V(2);
  V inv_j = V(1) / bar_jacobian(length, r);
  V det_j = bar_jacobian(length, r);
  V G = e.E / (V(2) * (V(1) + e.nu));
  V cs[6] = {e.A * e.E, G * e.A * e.k, G * e.A * e.k, G * e.it, e.E * e.i22p, e.E * e.i11p};
  // rows of the six terms (beam_local_k's bu .. btw): entries at disjoint positions
  std::vector<V> B(6 * n, V(0));
  const V d1 = std::fabs(bar_dh1(r) * inv_j), d2 = std::fabs(bar_dh2(r) * inv_j);
  const V h1 = std::fabs(bar_h1(r)), h2 = std::fabs(bar_h2(r));
  B[0 * n + 0] = d1; B[0 * n + 6] = d2;
  B[1 * n + 1] = d1; B[1 * n + 7] = d2; B[1 * n + 5] = h1; B[1 * n + 11] = h2;
  B[2 * n + 2] = d1; B[2 * n + 8] = d2; B[2 * n + 4] = h1; B[2 * n + 10] = h2;
  B[3 * n + 3] = d1; B[3 * n + 9] = d2;
  B[4 * n + 4] = d1; B[4 * n + 10] = d2;
  B[5 * n + 5] = d1; B[5 * n + 11] = d2;
  std::vector<V> k(n * n, V(0));
  for (int t = 0; t < 6; ++t) {
    const V scale = std::fabs(cs[t] * det_j * alpha);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) k[i * n + j] += (B[t * n + i] * B[t * n + j]) * scale;
  }
  return k;
}
// plate_local_k with |B|^T |C| |B| |w| per integration point and the drilling term
template <class V>
std::vector<V> plate_local_k_abs(const Model<V>& m, const Plate<V>& e) {
  const int n = 24;
  QuadLocal<V> q = quad_local(m, e);
  V ips[4][2] = {{e.ip_rs[0], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[1]}, {e.ip_rs[0], e.ip_rs[1]}};
  std::vector<V> k(n * n, V(0));
  for (int ip = 0; ip < 4; ++ip) {
    V det;
    std::vector<V> bm, bb, bs;
    plate_b(q, ips[ip][0], ips[ip][1], bm, bb, bs, det);
    V nu = e.nu;
    V cm_mul = e.E * e.t / (V(1) - nu * nu);
    V cb_mul = e.E * (e.t * e.t * e.t) / (V(12) * (V(1) - nu * nu));
    V cs_mul = e.E * e.t * e.k / (V(2) * (V(1) + nu));
    std::vector<V> base = {V(1), nu, V(0), nu, V(1), V(0), V(0), V(0), (V(1) - nu) / V(2)};
    std::vector<V> cm(9), cb(9);
    for (int i = 0; i < 9; ++i) { cm[i] = std::fabs(base[i] * cm_mul); cb[i] = std::fabs(base[i] * cb_mul); }
    std::vector<V> cs = {std::fabs(cs_mul), V(0), V(0), std::fabs(cs_mul)};
    auto term = [&](const std::vector<V>& B, const std::vector<V>& C, int rows) {
      auto Ba = abs_all(B);
      auto Bt = transpose(Ba, rows, n);
      return matmul(matmul(Bt, C, n, rows, rows), Ba, n, rows, n);
    };
    auto km = term(bm, cm, 3), kb = term(bb, cb, 3), ks = term(bs, cs, 2);
    const V w = std::fabs(det);
    for (int i = 0; i < n * n; ++i) k[i] += (km[i] + kb[i] + ks[i]) * w;
  }
  for (int i = 0; i < 4; ++i) k[(6 * i + 5) * n + 6 * i + 5] += V(1.0f);
  return k;
}

std::vector<uint8_t> plate_local_pattern() {
  std::vector<uint8_t> P(576, 0);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      for (int i : {0, 1})
        for (int j : {0, 1}) P[(6 * a + i) * 24 + 6 * b + j] = 1;
      for (int i : {2, 3, 4})
        for (int j : {2, 3, 4}) P[(6 * a + i) * 24 + 6 * b + j] = 1;
      if (a == b) P[(6 * a + 5) * 24 + 6 * b + 5] = 1;
    }
  return P;
}

// ---------------------------------------------------------------- scatter
// methods_for_{truss,beam,plate}_data_handle.rs: blocks of K_e into the global K
template <class V>
void scatter(Model<V>& m, const std::vector<V>& K, const std::vector<uint8_t>& pattern, int n,
             const std::vector<size_t>& node_index, int node_dof) {
  int nn = int(node_index.size());
  for (int a = 0; a < nn; ++a)
    for (int b = 0; b < nn; ++b)
      for (int i = 0; i < node_dof; ++i)
        for (int j = 0; j < node_dof; ++j) {
          int li = a * node_dof + i, lj = b * node_dof + j;
          V v = K[size_t(li) * n + lj];
          bool structural = pattern[size_t(li) * n + lj];
          if (v == V(0) && !structural) continue;
          m.t_row.push_back(int64_t(node_index[a]) * NODE_DOF + i);
          m.t_col.push_back(int64_t(node_index[b]) * NODE_DOF + j);
          m.t_val.push_back(v);
          m.t_structural.push_back(structural);
          m.t_nonzero.push_back(v != V(0));
          m.t_local.push_back(uint16_t(li * n + lj));
        }
}

// ---------------------------------------------------------------- add_truss
// src/fem/methods_for_truss_data_handle.rs:49-128, src/fem/structs/truss.rs:230-255
template <class V>
bool add_truss(Model<V>& m, uint32_t number, uint32_t n1, uint32_t n2, V E, V A, bool has_a2, V A2) {
  if (!check_node_exist(m, n1) || !check_node_exist(m, n2)) return false;
  if (m.trusses.count(number)) {
    m.last_error = "Truss element with number " + std::to_string(number) + " already exists!";
    return false;
  }
  uint64_t key = n1 < n2 ? (uint64_t(n1) << 32 | n2) : (uint64_t(n2) << 32 | n1);
  if (m.truss_pairs.count(key)) {
    m.last_error = "Truss element with node number " + std::to_string(n1) + " and " + std::to_string(n2) +
                   " already exists!";
    return false;
  }
  if (E <= V(0)) { m.last_error = "Young's modulus " + rust_debug(E) + " is less or equal to zero!"; return false; }
  if (A <= V(0)) { m.last_error = "Area " + rust_debug(A) + " is less or equal to zero!"; return false; }
  if (has_a2 && A2 <= V(0)) { m.last_error = "Area2 " + rust_debug(A2) + " is less or equal to zero!"; return false; }
  Truss<V> e{n1, n2, E, A, has_a2, A2, {}};
  auto d = element_vector(m, n1, n2);
  V L = norm3(d);
  std::string err;
  if (!rotation_to_align(d, Vec3<V>{L, V(0), V(0)}, m.abs_tol, e.R, err)) { m.last_error = err; return false; }
  auto k = truss_local_k(e, L);
  auto K = rotate_tkt(k, e.R, 6);
  auto P = rotate_pattern(truss_local_pattern(), e.R, 6);
  // only the translational 3x3 blocks are scattered (truss.rs TRUSS_NODE_DOF = 3)
  std::vector<size_t> idx = {m.nodes.at(n1).index, m.nodes.at(n2).index};
  const size_t t0 = m.t_row.size();
  scatter(m, K, P, 6, idx, 3);
  m.erange.push_back({0, number, t0, m.t_row.size()});
  m.trusses[number] = e;
  m.truss_pairs[key] = number;
  return true;
}

// src/fem/methods_for_beam_data_handle.rs:49-142, src/fem/structs/beam.rs:63-117,687-745
template <class V>
bool add_beam(Model<V>& m, uint32_t number, uint32_t n1, uint32_t n2, V E, V nu, V A, V i11, V i22, V i12, V it,
              V shear, Vec3<V> v1) {
  if (!check_node_exist(m, n1) || !check_node_exist(m, n2)) return false;
  if (m.beams.count(number)) {
    m.last_error = "Beam element with number " + std::to_string(number) + " already exists!";
    return false;
  }
  uint64_t key = n1 < n2 ? (uint64_t(n1) << 32 | n2) : (uint64_t(n2) << 32 | n1);
  if (m.beam_pairs.count(key)) {
    m.last_error = "Beam element with node number " + std::to_string(n1) + " and " + std::to_string(n2) +
                   " already exists!";
    return false;
  }
  auto msg = [&](const char* what, V v) {
    m.last_error = std::string(what) + " " + rust_debug(v) + " is less or equal to zero!";
    return false;
  };
  if (E <= V(0)) return msg("Young's modulus", E);
  if (nu <= V(0)) return msg("Poisson's ratio", E);  // the crate reports E here (beam.rs:84)
  if (A <= V(0)) return msg("Area", A);
  if (i11 <= V(0)) return msg("I11", i11);
  if (i22 <= V(0)) return msg("I22", i22);
  if (it <= V(0)) return msg("It", it);
  if (shear <= V(0)) return msg("Shear factor", shear);
  auto d = element_vector(m, n1, n2);
  V f = dot3(v1, d) / dot3(d, d);
  Vec3<V> p{v1[0] - f * d[0], v1[1] - f * d[1], v1[2] - f * d[2]};
  if (norm3(p) == V(0)) {
    m.last_error = "Local axis 1 direction [" + rust_debug(v1[0]) + ", " + rust_debug(v1[1]) + ", " +
                   rust_debug(v1[2]) + "] parallel to element " + std::to_string(number) + "!";
    return false;
  }
  Beam<V> e{n1, n2, E, nu, A, V(0), V(0), it, shear, {}};
  V angle;
  principal_moments(i11, i22, i12, m.rel_tol, e.i11p, e.i22p, angle);
  std::string err;
  if (!beam_rotation(d, v1, angle, m.rel_tol, m.abs_tol, e.R, err)) { m.last_error = err; return false; }
  V L = norm3(d);
  auto k = beam_local_k(e, L);
  auto K = rotate_tkt(k, e.R, 12);
  auto P = beam_global_pattern(e.R);
  std::vector<size_t> idx = {m.nodes.at(n1).index, m.nodes.at(n2).index};
  const size_t t0 = m.t_row.size();
  scatter(m, K, P, 12, idx, 6);
  m.erange.push_back({1, number, t0, m.t_row.size()});
  m.beams[number] = e;
  m.beam_pairs[key] = number;
  return true;
}

// src/fem/methods_for_plate_data_handle.rs:62-217, src/fem/structs/plate.rs:58-158,1021-1105
template <class V>
bool add_plate(Model<V>& m, uint32_t number, const uint32_t nn[4], V E, V nu, V t, V shear) {
  for (int i = 0; i < 4; ++i)
    if (!check_node_exist(m, nn[i])) return false;
  if (m.plates.count(number)) {
    m.last_error = "Plate element with number " + std::to_string(number) + " already exists!";
    return false;
  }
  // is_nodes_numbers_same: every given node belongs to an existing plate
  for (uint32_t cand : m.node_plates[nn[0]]) {
    const auto& pe = m.plates.at(cand);
    bool all = true;
    for (int i = 0; i < 4 && all; ++i)
      all = nn[i] == pe.n[0] || nn[i] == pe.n[1] || nn[i] == pe.n[2] || nn[i] == pe.n[3];
    if (all) {
      m.last_error = "Plate element with nodes numbers [" + std::to_string(nn[0]) + ", " + std::to_string(nn[1]) +
                     ", " + std::to_string(nn[2]) + ", " + std::to_string(nn[3]) + "] already exists!";
      return false;
    }
  }
  auto msg = [&](const char* what, V v) {
    m.last_error = std::string(what) + " " + rust_debug(v) + " is less or equal to zero!";
    return false;
  };
  if (E <= V(0)) return msg("Young's modulus", E);
  if (nu <= V(0)) return msg("Poisson's ratio", E);  // plate.rs:79 reports E
  if (t <= V(0)) return msg("Thickness", E);         // plate.rs:82 reports E
  if (shear <= V(0)) return msg("Shear factor", shear);
  Vec3<V> pts[4];
  for (int i = 0; i < 4; ++i) pts[i] = coords(m, nn[i]);
  if (quad_on_line(pts[0], pts[1], pts[2], pts[3])) {
    m.last_error = "Some nodes of " + std::to_string(number) + " element lie on the line!";
    return false;
  }
  if (!quad_on_plane(pts[0], pts[1], pts[2], pts[3], m.abs_tol)) {
    m.last_error = "Not all nodes of element " + std::to_string(number) + " lie on the plane!";
    return false;
  }
  size_t hull_len = 0;
  std::string err;
  if (!quad_convex(nn, pts, m.abs_tol, hull_len, err)) { m.last_error = err; return false; }
  if (hull_len != 4) {
    m.last_error = "Element " + std::to_string(number) + " non-convex!";
    return false;
  }
  Plate<V> e;
  for (int i = 0; i < 4; ++i) e.n[i] = nn[i];
  e.E = E; e.nu = nu; e.t = t; e.k = shear;
  if (!quad_rotation(pts[1], pts[2], pts[3], m.abs_tol, e.R, err)) { m.last_error = err; return false; }
  V g = std::sqrt(V(1.0f / 3.0f)) * V(1);  // f32-rounded 1/3, plate.rs:1068
  e.ip_rs = {g, std::sqrt(V(1.0f / 3.0f)) * V(-1), V(0), V(0)};
  auto k = plate_local_k(m, e);
  auto K = rotate_tkt(k, e.R, 24);
  auto P = rotate_pattern(plate_local_pattern(), e.R, 24);
  std::vector<size_t> idx = {m.nodes.at(nn[0]).index, m.nodes.at(nn[1]).index, m.nodes.at(nn[2]).index,
                             m.nodes.at(nn[3]).index};
  const size_t t0 = m.t_row.size();
  scatter(m, K, P, 24, idx, 6);
  m.erange.push_back({2, number, t0, m.t_row.size()});
  m.plates[number] = e;
  for (int i = 0; i < 4; ++i) m.node_plates[nn[i]].push_back(number);
  return true;
}

// ---------------------------------------------------------------- export
// mode 0: the crate's value-based matrix (entries added only when != 0, then
//         summed entries == 0 dropped as in methods_for_separate_stiffness_matrix.rs:278)
// mode 1: structural pattern (union of the elements' structural patterns), values summed
template <class V>
int64_t export_csr(Model<V>& m, int mode) {
  size_t nt = m.t_row.size();
  std::vector<size_t> order;
  order.reserve(nt);
  for (size_t i = 0; i < nt; ++i) {
    if (mode == 0 && !m.t_nonzero[i]) continue;
    if (mode == 1 && !m.t_structural[i]) continue;
    order.push_back(i);
  }
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return m.t_row[a] != m.t_row[b] ? m.t_row[a] < m.t_row[b] : m.t_col[a] < m.t_col[b];
  });
  int64_t ndof = int64_t(m.nodes_number) * NODE_DOF;
  m.row_ptr.assign(ndof + 1, 0);
  m.col_idx.clear();
  m.val.clear();
  size_t i = 0;
  while (i < order.size()) {
    size_t j = i;
    V acc = m.t_val[order[i]];
    while (j + 1 < order.size() && m.t_row[order[j + 1]] == m.t_row[order[i]] &&
           m.t_col[order[j + 1]] == m.t_col[order[i]]) {
      ++j;
      acc += m.t_val[order[j]];
    }
    if (!(mode == 0 && acc == V(0))) {
      m.row_ptr[m.t_row[order[i]] + 1] += 1;
      m.col_idx.push_back(int32_t(m.t_col[order[i]]));
      m.val.push_back(double(acc));
    }
    i = j + 1;
  }
  for (int64_t r = 0; r < ndof; ++r) m.row_ptr[r + 1] += m.row_ptr[r];
  return int64_t(m.val.size());
}

// Term magnitude of every entry of the structural CSR (export_csr mode 1: the
// same entries in the same order): the sum over the entry's contributions of
// the element magnitude matrices (|T|^T k^ |T|, k^ from *_local_k_abs).
template <class V>
void export_abs(const Model<V>& m, std::vector<double>& out) {
  std::vector<V> ta(m.t_val.size(), V(0));
  for (const auto& er : m.erange) {
    std::vector<V> K;
    if (er.type == 0) {
      const auto& e = m.trusses.at(er.number);
      K = rotate_tkt_abs(abs_all(truss_local_k(e, norm3(element_vector(m, e.n1, e.n2)))), e.R, 6);
    } else if (er.type == 1) {
      const auto& e = m.beams.at(er.number);
      K = rotate_tkt_abs(beam_local_k_abs(e, norm3(element_vector(m, e.n1, e.n2))), e.R, 12);
    } else {
      const auto& e = m.plates.at(er.number);
      K = rotate_tkt_abs(plate_local_k_abs(m, e), e.R, 24);
    }
    for (size_t t = er.t0; t < er.t1; ++t) ta[t] = K[m.t_local[t]];
  }
  std::vector<size_t> order;
  for (size_t i = 0; i < m.t_row.size(); ++i)
    if (m.t_structural[i]) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return m.t_row[a] != m.t_row[b] ? m.t_row[a] < m.t_row[b] : m.t_col[a] < m.t_col[b];
  });
  out.clear();
  size_t i = 0;
  while (i < order.size()) {
    size_t j = i;
    V acc = ta[order[i]];
    while (j + 1 < order.size() && m.t_row[order[j + 1]] == m.t_row[order[i]] &&
           m.t_col[order[j + 1]] == m.t_col[order[i]]) {
      ++j;
      acc += ta[order[j]];
    }
    out.push_back(double(acc));
    i = j + 1;
  }
}

// ---------------------------------------------------------------- boundary conditions
const char* const DOF_NAMES[6] = {"X", "Y", "Z", "ThX", "ThY", "ThZ"};  // DOFParameter Debug

template <class V>
void ensure_bc(Model<V>& m) {
  const size_t n = size_t(m.nodes_number) * NODE_DOF;
  if (m.constrained.size() != n) {
    m.constrained.assign(n, 0);
    m.displacement.assign(n, V(0));
    m.force.assign(n, V(0));
  }
}

// methods_for_bc_data_handle.rs:176-203
template <class V>
bool add_displacement(Model<V>& m, uint32_t number, int dof, V value) {
  if (!check_node_exist(m, number)) return false;
  ensure_bc(m);
  const size_t index = m.nodes.at(number).index * NODE_DOF + size_t(dof);
  if (m.constrained[index]) {
    m.last_error = std::string("Displacement ") + DOF_NAMES[dof] + " already applied to node " +
                   std::to_string(number) + "!";
    return false;
  }
  m.constrained[index] = 1;
  m.displacement[index] = value;
  return true;
}

// methods_for_bc_data_handle.rs:34-56
template <class V>
bool add_concentrated_load(Model<V>& m, uint32_t number, int dof, V value) {
  if (!check_node_exist(m, number)) return false;
  ensure_bc(m);
  const size_t index = m.nodes.at(number).index * NODE_DOF + size_t(dof);
  m.force[index] = m.force[index] + value;
  return true;
}

// methods_for_bc_data_handle.rs:58-100, beam.rs:775-797: q L / 2-type nodal
// loads from one integration point (r = 0, alpha = 2)
template <class V>
bool add_line_load(Model<V>& m, uint32_t number, int dof, V value) {
  auto it = m.beams.find(number);
  if (it == m.beams.end()) {
    m.last_error = "Beam element with number " + std::to_string(number) + " does not exist!";
    return false;
  }
  ensure_bc(m);
  const auto& e = it->second;
  const V r = V(0), alpha = V(2);
  const V det = bar_jacobian(norm3(element_vector(m, e.n1, e.n2)), r);
  const V f[2] = {V(0) + (bar_h1(r) * value) * (det * alpha), V(0) + (bar_h2(r) * value) * (det * alpha)};
  const size_t i1 = m.nodes.at(e.n1).index * NODE_DOF + size_t(dof), i2 = m.nodes.at(e.n2).index * NODE_DOF + size_t(dof);
  m.force[i1] = m.force[i1] + f[0];
  m.force[i2] = m.force[i2] + f[1];
  return true;
}

// methods_for_bc_data_handle.rs:102-174, plate.rs:1145-1185: nodal loads
// h_a(r, s) p det J(r, s) alpha_r alpha_s summed over the four points
template <class V>
bool add_surface_load(Model<V>& m, uint32_t number, int dof, V value) {
  auto it = m.plates.find(number);
  if (it == m.plates.end()) {
    m.last_error = "Plate element with number " + std::to_string(number) + " does not exist!";
    return false;
  }
  ensure_bc(m);
  const auto& e = it->second;
  const QuadLocal<V> q = quad_local(m, e);
  const V ips[4][2] = {{e.ip_rs[0], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[0]}, {e.ip_rs[1], e.ip_rs[1]},
                       {e.ip_rs[0], e.ip_rs[1]}};
  V f[4] = {V(0), V(0), V(0), V(0)};
  for (int ip = 0; ip < 4; ++ip) {
    const V r = ips[ip][0], s = ips[ip][1];
    V J[4];
    quad_jacobian(q, r, s, J);
    const V det = J[0] * J[3] - J[1] * J[2];
    const V qf = V(0.25f);
    const V h[4] = {qf * (V(1) + r) * (V(1) + s), qf * (V(1) - r) * (V(1) + s), qf * (V(1) - r) * (V(1) - s),
                    qf * (V(1) + r) * (V(1) - s)};
    for (int a = 0; a < 4; ++a) f[a] = f[a] + (h[a] * value) * (det * (V(1) * V(1)));
  }
  for (int a = 0; a < 4; ++a) {
    const size_t ia = m.nodes.at(e.n[a]).index * NODE_DOF + size_t(dof);
    m.force[ia] = m.force[ia] + f[a];
  }
  return true;
}

// methods_for_separate_stiffness_matrix.rs:217-320 on the exported CSR (call
// export_csr first).  Entries are classified exactly as the crate does; the
// triplets come out sorted by (row, column) instead of in HashMap order.
template <class V>
bool separate_sparse(Model<V>& m) {
  ensure_bc(m);
  const int64_t n = int64_t(m.nodes_number) * NODE_DOF;
  std::vector<double> diag(size_t(n), 0.0);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t k = m.row_ptr[r]; k < m.row_ptr[r + 1]; ++k)
      if (m.col_idx[k] == r) diag[size_t(r)] = m.val[k];
  m.aa_idx.clear();
  m.bb_idx.clear();
  for (int64_t idx = 0; idx < n; ++idx) {
    if (diag[size_t(idx)] == 0.0) {
      if (m.constrained[size_t(idx)]) {
        uint32_t number = 0;
        for (auto& kv : m.nodes)
          if (kv.second.index == size_t(idx / NODE_DOF)) number = kv.first;
        m.last_error = std::string("There are no stiffness to withstand displacement ") + DOF_NAMES[idx % NODE_DOF] +
                       " applied to node " + std::to_string(number) + "!";
        return false;
      }
      continue;  // excluded / inactive dof
    }
    (m.constrained[size_t(idx)] ? m.bb_idx : m.aa_idx).push_back(idx);
  }
  if (m.bb_idx.empty()) {
    m.last_error = "No restraints";
    return false;
  }
  std::vector<int64_t> aa_pos(size_t(n), -1), bb_pos(size_t(n), -1);
  for (size_t i = 0; i < m.aa_idx.size(); ++i) aa_pos[size_t(m.aa_idx[i])] = int64_t(i);
  for (size_t i = 0; i < m.bb_idx.size(); ++i) bb_pos[size_t(m.bb_idx[i])] = int64_t(i);
  const int64_t rows[4] = {int64_t(m.aa_idx.size()), int64_t(m.aa_idx.size()), int64_t(m.bb_idx.size()),
                           int64_t(m.bb_idx.size())};
  for (int w = 0; w < 4; ++w) {
    m.sep_rp[w].assign(size_t(rows[w]) + 1, 0);
    m.sep_ci[w].clear();
    m.sep_v[w].clear();
  }
  for (int64_t r = 0; r < n; ++r) {
    const int64_t ia = aa_pos[size_t(r)], ib = bb_pos[size_t(r)];
    for (int64_t k = m.row_ptr[r]; k < m.row_ptr[r + 1]; ++k) {
      const double v = m.val[k];
      if (v == 0.0) continue;
      const int64_t c = m.col_idx[k], ja = aa_pos[size_t(c)], jb = bb_pos[size_t(c)];
      int w = -1;
      int64_t row = 0, col = 0;
      if (ia >= 0 && ja >= 0) { w = 0; row = ia; col = ja; }
      else if (ia >= 0 && jb >= 0) { w = 1; row = ia; col = jb; }
      else if (ib >= 0 && ja >= 0) { w = 2; row = ib; col = ja; }
      else if (ib >= 0 && jb >= 0) { w = 3; row = ib; col = jb; }
      if (w < 0) continue;  // entry touches an excluded dof
      m.sep_rp[w][size_t(row) + 1] += 1;
      m.sep_ci[w].push_back(int32_t(col));
      m.sep_v[w].push_back(v);
    }
  }
  for (int w = 0; w < 4; ++w)
    for (int64_t r = 0; r < rows[w]; ++r) m.sep_rp[w][size_t(r) + 1] += m.sep_rp[w][size_t(r)];
  if (m.sep_v[0].empty()) {
    m.last_error = "Sparse separation: K_aa is empty (structure has no free stiffness?)";
    return false;
  }
  return true;
}

// ---------------------------------------------------------------- element analysis
// local = T g with T = block-diagonal copies of R (compose_rotation_matrix),
// as the crate's Matrix::multiply: a full row-by-column sum
template <class V>
std::vector<V> to_local(const std::array<V, 9>& R, const std::vector<V>& g) {
  const int n = int(g.size());
  std::vector<V> out(size_t(n), V(0));
  for (int i = 0; i < n; ++i) {
    V acc = V(0);
    for (int k = 0; k < n; ++k) {
      const V t = (i / 3 == k / 3) ? R[size_t((i % 3) * 3 + k % 3)] : V(0);
      acc += t * g[size_t(k)];
    }
    out[size_t(i)] = acc;
  }
  return out;
}
template <class V>
V row_dot(const std::vector<V>& b, const std::vector<V>& x, int off = 0) {
  V acc = V(0);
  for (size_t k = 0; k < x.size(); ++k) acc += b[size_t(off) + k] * x[k];
  return acc;
}

// truss.rs:281-333: ForceR
template <class V>
std::vector<V> truss_forces(const Model<V>& m, const Truss<V>& e, const std::vector<V>& u) {
  const size_t i1 = m.nodes.at(e.n1).index, i2 = m.nodes.at(e.n2).index;
  std::vector<V> g(6);
  for (int i = 0; i < 3; ++i) {
    g[size_t(i)] = u[i1 * NODE_DOF + size_t(i)];
    g[size_t(3 + i)] = u[i2 * NODE_DOF + size_t(i)];
  }
  auto loc = to_local(e.R, g);
  const V L = norm3(element_vector(m, e.n1, e.n2));
  const V r = V(0);
  const V inv_j = V(1) / bar_jacobian(L, r);
  std::vector<V> B = {V(0) + bar_dh1(r) * inv_j, V(0), V(0), V(0) + bar_dh2(r) * inv_j, V(0), V(0)};
  const V area = V(0) + (e.has_a2 ? (e.A2 - e.A) / V(2) * r + e.A - (e.A2 - e.A) / V(2) * V(-1) : e.A);
  const V strain = row_dot(B, loc);
  return {strain * (e.E * area / V(1))};
}

// beam.rs:803-993: ForceR, ForceS, ForceT, MomentR, MomentS (node 1, average,
// node 2), MomentT (node 1, average, node 2)
template <class V>
std::vector<V> beam_forces(const Model<V>& m, const Beam<V>& e, const std::vector<V>& u) {
  const size_t i1 = m.nodes.at(e.n1).index, i2 = m.nodes.at(e.n2).index;
  std::vector<V> g(12);
  for (int i = 0; i < 6; ++i) {
    g[size_t(i)] = u[i1 * NODE_DOF + size_t(i)];
    g[size_t(6 + i)] = u[i2 * NODE_DOF + size_t(i)];
  }
  auto loc = to_local(e.R, g);
  const V L = norm3(element_vector(m, e.n1, e.n2));
  const V r = V(0);
  const V inv_j = V(1) / bar_jacobian(L, r);
  auto row = [&](int a, int b, int c = -1, int d = -1) {
    std::vector<V> x(12, V(0));
    x[size_t(a)] = bar_dh1(r) * inv_j;
    x[size_t(b)] = bar_dh2(r) * inv_j;
    if (c >= 0) {
      x[size_t(c)] = x[size_t(c)] - bar_h1(r);
      x[size_t(d)] = x[size_t(d)] - bar_h2(r);
    }
    for (auto& v : x) v = V(0) + v;  // summed over the single point
    return x;
  };
  const V G = e.E / (V(2) * (V(1) + e.nu));
  const V fr = row_dot(row(0, 6), loc) * (e.E * e.A / V(1));
  const V fs = row_dot(row(1, 7, 5, 11), loc) * (G * e.A * e.k / V(1));
  const V ft = row_dot(row(2, 8, 4, 10), loc) * (G * e.A * e.k / V(1));
  const V mr = row_dot(row(3, 9), loc) * (G * e.it / V(1));
  const V ms = row_dot(row(4, 10), loc) * (e.E * e.i22p / V(1));
  const V mt = row_dot(row(5, 11), loc) * (e.E * e.i11p / V(1));
  return {fr, fs, ft, mr, ms + L * ft / V(2), ms, ms - L * ft / V(2), mt + L * fs / V(2), mt, mt - L * fs / V(2)};
}

// plate.rs:1196-1409: membrane forces R, S, RS, bending moments R, S, RS,
// shear forces RT, ST from B summed over the four nodal points
template <class V>
std::vector<V> plate_forces(const Model<V>& m, const Plate<V>& e, const std::vector<V>& u) {
  std::vector<V> g(24);
  for (int a = 0; a < 4; ++a) {
    const size_t ia = m.nodes.at(e.n[a]).index;
    for (int i = 0; i < 6; ++i) g[size_t(6 * a + i)] = u[ia * NODE_DOF + size_t(i)];
  }
  auto loc = to_local(e.R, g);
  const QuadLocal<V> q = quad_local(m, e);
  const int n = 24;
  std::vector<V> sm(3 * n, V(0)), sb(3 * n, V(0)), ss(2 * n, V(0));
  const V pts[4][2] = {{V(1), V(1)}, {V(-1), V(1)}, {V(-1), V(-1)}, {V(1), V(-1)}};
  for (int k = 0; k < 4; ++k) {
    std::vector<V> bm, bb, bs;
    V det;
    plate_b(q, pts[k][0], pts[k][1], bm, bb, bs, det);
    for (int i = 0; i < 3 * n; ++i) { sm[size_t(i)] = sm[size_t(i)] + bm[size_t(i)]; sb[size_t(i)] = sb[size_t(i)] + bb[size_t(i)]; }
    for (int i = 0; i < 2 * n; ++i) ss[size_t(i)] = ss[size_t(i)] + bs[size_t(i)];
  }
  const V nu = e.nu;
  const V base[9] = {V(1), nu, V(0), nu, V(1), V(0), V(0), V(0), (V(1) - nu) / V(2)};
  const V cmm = e.E / (V(1) - nu * nu);
  const V cbm = e.E * e.t / (V(2) * (V(1) - nu * nu));
  const V csm = e.E / (V(2) * (V(1) + nu));
  V em[3], eb[3], es[2];
  for (int i = 0; i < 3; ++i) { em[i] = row_dot(sm, loc, i * n); eb[i] = row_dot(sb, loc, i * n); }
  for (int i = 0; i < 2; ++i) es[i] = row_dot(ss, loc, i * n);
  V fm[3], fb[3], fsh[2];
  for (int i = 0; i < 3; ++i) {
    V am = V(0), ab = V(0);
    for (int j = 0; j < 3; ++j) { am += (base[i * 3 + j] * cmm) * em[j]; ab += (base[i * 3 + j] * cbm) * eb[j]; }
    fm[i] = am * (e.t / V(4));
    fb[i] = ab * (e.t * e.t / V(24));
  }
  for (int i = 0; i < 2; ++i) {
    V a = V(0);
    for (int j = 0; j < 2; ++j) a += ((i == j ? V(1) : V(0)) * csm) * es[j];
    fsh[i] = a * (e.t * e.k / V(4));
  }
  return {fm[0], fm[1], fm[2], fb[1], fb[0], fb[2], fsh[0], fsh[1]};
}

// ---------------------------------------------------------------- timed CPU baseline
// Steady-state value recompute of the assembled K on the host, for the CPU
// baseline: every element matrix again (dense local k, (T^T k) T, as add_*
// computes it), then every CSR value as the sum of its contributions in
// element-insertion order (the order export_csr sums them in).  Elements are
// spread over OpenMP threads; each CSR value is summed by one thread, so the
// result does not depend on the thread count.  Returns the median seconds of
// `reps` runs and the number of values that differ from export_csr(mode 1).
template <class V>
double time_values(Model<V>& m, int threads, int warmup, int reps, int64_t* mismatches) {
  // contributions of each structural CSR value, in insertion order (as export_csr mode 1)
  std::vector<size_t> order;
  order.reserve(m.t_row.size());
  for (size_t i = 0; i < m.t_row.size(); ++i)
    if (m.t_structural[i]) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return m.t_row[a] != m.t_row[b] ? m.t_row[a] < m.t_row[b] : m.t_col[a] < m.t_col[b];
  });
  std::vector<size_t> slot;  // slot s sums order[slot[s] .. slot[s + 1])
  for (size_t i = 0; i < order.size(); ++i)
    if (i == 0 || m.t_row[order[i]] != m.t_row[order[i - 1]] || m.t_col[order[i]] != m.t_col[order[i - 1]])
      slot.push_back(i);
  slot.push_back(order.size());
  const size_t ns = slot.size() - 1;
  std::vector<V> tv(m.t_val.size());
  std::vector<double> val(ns);
  std::vector<double> times;
  for (int rep = -warmup; rep < reps; ++rep) {  // negative reps: untimed warm-up
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
    for (int64_t ei = 0; ei < int64_t(m.erange.size()); ++ei) {
      const auto& er = m.erange[size_t(ei)];
      std::vector<V> K;
      int n = 0;
      if (er.type == 0) {
        const auto& e = m.trusses.at(er.number);
        auto d = element_vector(m, e.n1, e.n2);
        K = rotate_tkt(truss_local_k(e, norm3(d)), e.R, 6);
        n = 6;
      } else if (er.type == 1) {
        const auto& e = m.beams.at(er.number);
        auto d = element_vector(m, e.n1, e.n2);
        K = rotate_tkt(beam_local_k(e, norm3(d)), e.R, 12);
        n = 12;
      } else {
        const auto& e = m.plates.at(er.number);
        K = rotate_tkt(plate_local_k(m, e), e.R, 24);
        n = 24;
      }
      (void)n;
      for (size_t t = er.t0; t < er.t1; ++t) tv[t] = K[m.t_local[t]];
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t s = 0; s < int64_t(ns); ++s) {
      V acc = tv[order[slot[size_t(s)]]];
      for (size_t j = slot[size_t(s)] + 1; j < slot[size_t(s) + 1]; ++j) acc += tv[order[j]];
      val[size_t(s)] = double(acc);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (rep >= 0) times.push_back(std::chrono::duration<double>(t1 - t0).count());
  }
  std::sort(times.begin(), times.end());
  int64_t bad = 0;
  export_csr(m, 1);
  if (m.val.size() != ns) bad = -1;
  else
    for (size_t i = 0; i < ns; ++i) bad += std::memcmp(&m.val[i], &val[i], sizeof(double)) != 0;
  if (mismatches) *mismatches = bad;
  return times.empty() ? 0.0 : times[times.size() / 2];
}

struct AnyModel {
  bool f32;
  Model<double>* d;
  Model<float>* f;
};

}  // namespace

// ======================================================================= C ABI
extern "C" {

AnyModel* om_create(double rel_tol, double abs_tol, uint32_t nodes_number, int use_f32) {
  auto* am = new AnyModel{use_f32 != 0, nullptr, nullptr};
  if (am->f32) am->f = new Model<float>{float(rel_tol), float(abs_tol), nodes_number};
  else am->d = new Model<double>{rel_tol, abs_tol, nodes_number};
  return am;
}
void om_destroy(AnyModel* am) {
  delete am->d;
  delete am->f;
  delete am;
}
const char* om_last_error(AnyModel* am) { return am->f32 ? am->f->last_error.c_str() : am->d->last_error.c_str(); }

int64_t om_add_nodes(AnyModel* am, int64_t n, const uint32_t* numbers, const double* xyz) {
  for (int64_t i = 0; i < n; ++i) {
    bool ok = am->f32 ? add_node<float>(*am->f, numbers[i], float(xyz[3 * i]), float(xyz[3 * i + 1]), float(xyz[3 * i + 2]))
                      : add_node<double>(*am->d, numbers[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (!ok) return i;
  }
  return n;
}
// A2 entries that are NaN mean `None`
int64_t om_add_trusses(AnyModel* am, int64_t n, const uint32_t* numbers, const uint32_t* conn, const double* E,
                       const double* A, const double* A2) {
  for (int64_t i = 0; i < n; ++i) {
    bool has = !std::isnan(A2[i]);
    bool ok = am->f32 ? add_truss<float>(*am->f, numbers[i], conn[2 * i], conn[2 * i + 1], float(E[i]), float(A[i]), has, float(A2[i]))
                      : add_truss<double>(*am->d, numbers[i], conn[2 * i], conn[2 * i + 1], E[i], A[i], has, A2[i]);
    if (!ok) return i;
  }
  return n;
}
// props: [E, nu, A, i11, i22, i12, it, shear, v1x, v1y, v1z] per beam
int64_t om_add_beams(AnyModel* am, int64_t n, const uint32_t* numbers, const uint32_t* conn, const double* props) {
  for (int64_t i = 0; i < n; ++i) {
    const double* p = props + 11 * i;
    bool ok;
    if (am->f32)
      ok = add_beam<float>(*am->f, numbers[i], conn[2 * i], conn[2 * i + 1], float(p[0]), float(p[1]), float(p[2]),
                           float(p[3]), float(p[4]), float(p[5]), float(p[6]), float(p[7]),
                           Vec3<float>{float(p[8]), float(p[9]), float(p[10])});
    else
      ok = add_beam<double>(*am->d, numbers[i], conn[2 * i], conn[2 * i + 1], p[0], p[1], p[2], p[3], p[4], p[5], p[6],
                            p[7], Vec3<double>{p[8], p[9], p[10]});
    if (!ok) return i;
  }
  return n;
}
// props: [E, nu, t, shear] per plate
int64_t om_add_plates(AnyModel* am, int64_t n, const uint32_t* numbers, const uint32_t* conn, const double* props) {
  for (int64_t i = 0; i < n; ++i) {
    const double* p = props + 4 * i;
    bool ok = am->f32 ? add_plate<float>(*am->f, numbers[i], conn + 4 * i, float(p[0]), float(p[1]), float(p[2]), float(p[3]))
                      : add_plate<double>(*am->d, numbers[i], conn + 4 * i, p[0], p[1], p[2], p[3]);
    if (!ok) return i;
  }
  return n;
}

// type: 0 truss, 1 beam, 2 plate
int om_get_rotation(AnyModel* am, int type, uint32_t number, double out[9]) {
  auto get = [&](auto& m) -> int {
    const char* names[3] = {"Truss", "Beam", "Plate"};
    bool found = type == 0 ? m.trusses.count(number) : type == 1 ? m.beams.count(number) : m.plates.count(number);
    if (!found) {
      m.last_error = std::string(names[type]) + " element with number " + std::to_string(number) + " does not exist!";
      return 1;
    }
    const auto& R = type == 0 ? m.trusses.at(number).R : type == 1 ? m.beams.at(number).R : m.plates.at(number).R;
    for (int i = 0; i < 9; ++i) out[i] = double(R[i]);
    return 0;
  };
  return am->f32 ? get(*am->f) : get(*am->d);
}

// element matrix in the local frame (frame=0) or after (T^T k) T (frame=1);
// writes n*n doubles, returns n (6, 12 or 24) or -1
int om_element_matrix(AnyModel* am, int type, uint32_t number, int frame, double* out) {
  auto get = [&](auto& m) -> int {
    using V = std::decay_t<decltype(m.rel_tol)>;
    std::vector<V> k;
    std::array<V, 9> R;
    int n;
    if (type == 0) {
      if (!m.trusses.count(number)) return -1;
      const auto& e = m.trusses.at(number);
      k = truss_local_k(e, norm3(element_vector(m, e.n1, e.n2)));
      R = e.R; n = 6;
    } else if (type == 1) {
      if (!m.beams.count(number)) return -1;
      const auto& e = m.beams.at(number);
      k = beam_local_k(e, norm3(element_vector(m, e.n1, e.n2)));
      R = e.R; n = 12;
    } else {
      if (!m.plates.count(number)) return -1;
      const auto& e = m.plates.at(number);
      k = plate_local_k(m, e);
      R = e.R; n = 24;
    }
    if (frame == 1) k = rotate_tkt(k, R, n);
    for (int i = 0; i < n * n; ++i) out[i] = double(k[i]);
    return n;
  };
  return am->f32 ? get(*am->f) : get(*am->d);
}

// Global-frame K_e of one element from explicit node coordinates (nodes_per x 3)
// and parameters (the fem_add_* layouts), computed as add_* does after its
// validation (no validation here).  `parts` selects stiffness terms (see
// plate_local_k / beam_local_k; ignored for trusses).  Writes n*n doubles and
// returns n (6, 12, 24), or -1 when no frame can be formed.  Used by the tests'
// reference gradients (exact part splits and finite differences).
int om_eval_element(int type, const double* xyz, const double* params, double rel_tol, double abs_tol, unsigned parts,
                    double* out) {
  const int nn = type == 2 ? 4 : 2;
  Model<double> m{rel_tol, abs_tol, uint32_t(nn)};
  for (int a = 0; a < nn; ++a)
    if (!add_node<double>(m, uint32_t(a + 1), xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2])) return -1;
  std::vector<double> K;
  std::string err;
  int n;
  if (type == 0) {
    Truss<double> e{1, 2, params[0], params[1], !std::isnan(params[2]), params[2], {}};
    auto d = element_vector(m, 1, 2);
    double L = norm3(d);
    if (!rotation_to_align(d, Vec3<double>{L, 0.0, 0.0}, abs_tol, e.R, err)) return -1;
    n = 6;
    K = rotate_tkt(truss_local_k(e, L), e.R, n);
  } else if (type == 1) {
    const double* p = params;
    Beam<double> e{1, 2, p[0], p[1], p[2], 0.0, 0.0, p[6], p[7], {}};
    double angle;
    principal_moments(p[3], p[4], p[5], rel_tol, e.i11p, e.i22p, angle);
    auto d = element_vector(m, 1, 2);
    if (!beam_rotation(d, Vec3<double>{p[8], p[9], p[10]}, angle, rel_tol, abs_tol, e.R, err)) return -1;
    n = 12;
    K = rotate_tkt(beam_local_k(e, norm3(d), parts), e.R, n);
  } else {
    Plate<double> e;
    for (int i = 0; i < 4; ++i) e.n[i] = uint32_t(i + 1);
    e.E = params[0]; e.nu = params[1]; e.t = params[2]; e.k = params[3];
    if (!quad_rotation(coords(m, 2), coords(m, 3), coords(m, 4), abs_tol, e.R, err)) return -1;
    double g = std::sqrt(double(1.0f / 3.0f)) * 1.0;
    e.ip_rs = {g, std::sqrt(double(1.0f / 3.0f)) * -1.0, 0.0, 0.0};
    n = 24;
    K = rotate_tkt(plate_local_k(m, e, parts), e.R, n);
  }
  for (int i = 0; i < n * n; ++i) out[i] = K[i];
  return n;
}

// principal moments of inertia (beam.rs:119-160): out = [i11_p, i22_p, angle]
// quadrilateral predicates re-exported by the crate (src/lib.rs:19-23); points: 4 x 3
int om_quad_on_line(const double* p) {
  Vec3<double> v[4];
  for (int k = 0; k < 4; ++k) v[k] = Vec3<double>{p[3 * k], p[3 * k + 1], p[3 * k + 2]};
  return quad_on_line(v[0], v[1], v[2], v[3]) ? 1 : 0;
}
int om_quad_on_plane(const double* p, double abs_tol) {
  Vec3<double> v[4];
  for (int k = 0; k < 4; ++k) v[k] = Vec3<double>{p[3 * k], p[3 * k + 1], p[3 * k + 2]};
  return quad_on_plane(v[0], v[1], v[2], v[3], abs_tol) ? 1 : 0;
}
// hull point numbers into out (up to 4); returns their count, -1 without a frame
int om_convex_hull(const uint32_t* num, const double* p, double abs_tol, uint32_t* out) {
  Vec3<double> v[4];
  for (int k = 0; k < 4; ++k) v[k] = Vec3<double>{p[3 * k], p[3 * k + 1], p[3 * k + 2]};
  std::vector<HullPoint<double>> h;
  std::string err;
  if (!quad_hull(num, v, abs_tol, h, err)) return -1;
  for (size_t i = 0; i < h.size(); ++i) out[i] = h[i].n;
  return int(h.size());
}

void om_principal_moments(double i11, double i22, double i12, double rel_tol, double out[3]) {
  principal_moments<double>(i11, i22, i12, rel_tol, out[0], out[1], out[2]);
}

// Rust `{:?}` of a float, for tests of the message formatter
int om_rust_debug(double x, int use_f32, char* buf, int len) {
  std::string s = use_f32 ? rust_debug<float>(float(x)) : rust_debug<double>(x);
  std::snprintf(buf, len, "%s", s.c_str());
  return int(s.size());
}

int64_t om_export(AnyModel* am, int mode) { return am->f32 ? export_csr(*am->f, mode) : export_csr(*am->d, mode); }
int64_t om_ndof(AnyModel* am) { return int64_t(am->f32 ? am->f->nodes_number : am->d->nodes_number) * NODE_DOF; }
void om_export_copy(AnyModel* am, int64_t* row_ptr, int32_t* col_idx, double* val) {
  auto cp = [&](auto& m) {
    std::memcpy(row_ptr, m.row_ptr.data(), m.row_ptr.size() * sizeof(int64_t));
    std::memcpy(col_idx, m.col_idx.data(), m.col_idx.size() * sizeof(int32_t));
    std::memcpy(val, m.val.data(), m.val.size() * sizeof(double));
  };
  if (am->f32) cp(*am->f); else cp(*am->d);
}
// term magnitudes of the structural CSR's entries (om_export(., 1) order);
// returns the number of entries written
int64_t om_export_magnitude(AnyModel* am, double* out, int64_t cap) {
  std::vector<double> v;
  if (am->f32) export_abs(*am->f, v); else export_abs(*am->d, v);
  if (int64_t(v.size()) > cap) return -1;
  std::memcpy(out, v.data(), v.size() * sizeof(double));
  return int64_t(v.size());
}
int64_t om_num_contributions(AnyModel* am) { return int64_t(am->f32 ? am->f->t_val.size() : am->d->t_val.size()); }
// largest |value| among non-structural contributions, relative to the largest |value| overall
double om_nonstructural_max_rel(AnyModel* am) {
  auto f = [](auto& m) {
    double mx = 0, ns = 0;
    for (size_t i = 0; i < m.t_val.size(); ++i) {
      double v = std::fabs(double(m.t_val[i]));
      mx = std::max(mx, v);
      if (!m.t_structural[i]) ns = std::max(ns, v);
    }
    return mx > 0 ? ns / mx : 0.0;
  };
  return am->f32 ? f(*am->f) : f(*am->d);
}
// contributions that are non-zero although the structural rule says zero
int64_t om_structural_violations(AnyModel* am) {
  auto cnt = [](auto& m) {
    int64_t c = 0;
    for (size_t i = 0; i < m.t_val.size(); ++i) c += m.t_nonzero[i] && !m.t_structural[i];
    return c;
  };
  return am->f32 ? cnt(*am->f) : cnt(*am->d);
}

int64_t om_num_elements(AnyModel* am, int type) {
  auto cnt = [&](auto& m) {
    int64_t c = 0;
    for (const auto& er : m.erange) c += er.type == type;
    return c;
  };
  return am->f32 ? cnt(*am->f) : cnt(*am->d);
}

// element forces from global displacements u (6 x nodes_number) for the
// elements of `type` in insertion order (truss 1, beam 10, plate 8 values each)
int64_t om_element_forces(AnyModel* am, int type, const double* u, double* out) {
  auto go = [&](auto& m, auto tag) -> int64_t {
    using V = decltype(tag);
    std::vector<V> uu(size_t(m.nodes_number) * NODE_DOF);
    for (size_t i = 0; i < uu.size(); ++i) uu[i] = V(u[i]);
    int64_t w = 0;
    for (const auto& er : m.erange) {
      if (er.type != type) continue;
      std::vector<V> f = type == 0 ? truss_forces(m, m.trusses.at(er.number), uu)
                         : type == 1 ? beam_forces(m, m.beams.at(er.number), uu)
                                     : plate_forces(m, m.plates.at(er.number), uu);
      for (V x : f) out[w++] = double(x);
    }
    return w;
  };
  return am->f32 ? go(*am->f, float(0)) : go(*am->d, double(0));
}

// CPU baseline: median seconds of `reps` value recomputes on `threads` OpenMP
// threads; *mismatches = values differing from the sequential export.
double om_time_values(AnyModel* am, int threads, int warmup, int reps, int64_t* mismatches) {
  return am->f32 ? time_values(*am->f, threads, warmup, reps, mismatches)
                 : time_values(*am->d, threads, warmup, reps, mismatches);
}

// boundary conditions; dof 0..5 = X, Y, Z, ThX, ThY, ThZ.  Return 0 or 1 (message in om_last_error).
int om_add_displacement(AnyModel* am, uint32_t number, int dof, double value) {
  bool ok = am->f32 ? add_displacement<float>(*am->f, number, dof, float(value))
                    : add_displacement<double>(*am->d, number, dof, value);
  return ok ? 0 : 1;
}
int om_add_concentrated_load(AnyModel* am, uint32_t number, int dof, double value) {
  bool ok = am->f32 ? add_concentrated_load<float>(*am->f, number, dof, float(value))
                    : add_concentrated_load<double>(*am->d, number, dof, value);
  return ok ? 0 : 1;
}
int om_add_line_load(AnyModel* am, uint32_t number, int dof, double value) {
  bool ok = am->f32 ? add_line_load<float>(*am->f, number, dof, float(value)) : add_line_load<double>(*am->d, number, dof, value);
  return ok ? 0 : 1;
}
int om_add_surface_load(AnyModel* am, uint32_t number, int dof, double value) {
  bool ok = am->f32 ? add_surface_load<float>(*am->f, number, dof, float(value))
                    : add_surface_load<double>(*am->d, number, dof, value);
  return ok ? 0 : 1;
}
// the force vector (6 x nodes_number)
void om_forces(AnyModel* am, double* out) {
  auto cp = [&](auto& m) {
    ensure_bc(m);
    for (size_t i = 0; i < m.force.size(); ++i) out[i] = double(m.force[i]);
  };
  if (am->f32) cp(*am->f); else cp(*am->d);
}
// separation of the exported CSR (om_export first); out[6] = n_aa, n_bb, nnz of aa, ab, ba, bb
int om_separate(AnyModel* am, int64_t out[6]) {
  auto run = [&](auto& m) {
    if (!separate_sparse(m)) return 1;
    out[0] = int64_t(m.aa_idx.size());
    out[1] = int64_t(m.bb_idx.size());
    for (int w = 0; w < 4; ++w) out[2 + w] = int64_t(m.sep_v[w].size());
    return 0;
  };
  return am->f32 ? run(*am->f) : run(*am->d);
}
void om_separated_copy(AnyModel* am, int which, int64_t* rp, int32_t* ci, double* v) {
  auto cp = [&](auto& m) {
    std::memcpy(rp, m.sep_rp[which].data(), m.sep_rp[which].size() * sizeof(int64_t));
    std::memcpy(ci, m.sep_ci[which].data(), m.sep_ci[which].size() * sizeof(int32_t));
    std::memcpy(v, m.sep_v[which].data(), m.sep_v[which].size() * sizeof(double));
  };
  if (am->f32) cp(*am->f); else cp(*am->d);
}
// k_aa_indexes, k_bb_indexes, r_a (compose_r_a_vector :323-333), u_b (compose_u_b_vector :335-345)
void om_separated_vectors(AnyModel* am, int64_t* aa, int64_t* bb, double* r_a, double* u_b) {
  auto cp = [&](auto& m) {
    for (size_t i = 0; i < m.aa_idx.size(); ++i) {
      aa[i] = m.aa_idx[i];
      r_a[i] = double(m.force[size_t(m.aa_idx[i])]);
    }
    for (size_t i = 0; i < m.bb_idx.size(); ++i) {
      bb[i] = m.bb_idx[i];
      u_b[i] = double(m.displacement[size_t(m.bb_idx[i])]);
    }
  };
  if (am->f32) cp(*am->f); else cp(*am->d);
}

}  // extern "C"
