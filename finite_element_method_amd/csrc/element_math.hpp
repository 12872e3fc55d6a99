// element_math.hpp -- device (gfx950) element kinematics and stiffness blocks.
//
// All arithmetic is FP64.  Each function cites the crate file it computes the
// same quantity as; the formulas are evaluated in closed form per node-pair
// block instead of the crate's dense map-backed products, so results agree with
// the crate to rounding (the parity bar is 1e-12 relative).
#pragma once
#include <hip/hip_runtime.h>

namespace fem {

constexpr int NODE_DOF = 6;
constexpr double PI_F32 = 3.1415927410125732;  // std::f32::consts::PI widened (beam.rs:7,149)
// sqrt(f64(1f32/3f32)): the crate's Gauss abscissa (plate.rs:1066-1091)
constexpr float ONE_THIRD_F32 = 1.0f / 3.0f;
constexpr double KROT6 = 1.0;  // plate.rs:25

// Frames (rotation matrices, the beam roll) are computed without contracting
// a * b + c into fused multiply-adds, operation for operation as the crate
// writes them: the beam roll is acos of a cosine that is often close to +-1,
// where a last-bit difference in the cosine moves the angle by ~1e-8.
#pragma clang fp contract(off)
__host__ __device__ inline double cwt(double v, double abs_tol) {  // math_functions.rs:3-12
  return fabs(v) < abs_tol ? 0.0 : v;
}

__host__ __device__ inline double norm3(double x, double y, double z) { return sqrt((x * x + y * y) + z * z); }

// Rotation R with R.(a/|a|) = b/|b| (minimal rotation about a x b), entries
// snapped with abs_tol; restates extended_matrix's
// Vector3::rotation_matrix_to_align_with_vector (call sites truss.rs:80-84,
// beam.rs:179-183, quadrilateral_4n_element_functions.rs:158-162).
// Returns false for a zero-length vector.
__host__ __device__ inline bool rotation_to_align(const double a[3], const double b[3], double abs_tol,
                                                  double R[9]) {
  double na = norm3(a[0], a[1], a[2]), nb = norm3(b[0], b[1], b[2]);
  if (na == 0.0 || nb == 0.0) return false;
  double ua[3] = {a[0] / na, a[1] / na, a[2] / na};
  double ub[3] = {b[0] / nb, b[1] / nb, b[2] / nb};
  double v[3] = {ua[1] * ub[2] - ua[2] * ub[1], ua[2] * ub[0] - ua[0] * ub[2], ua[0] * ub[1] - ua[1] * ub[0]};
  double s = norm3(v[0], v[1], v[2]);
  double c = (ua[0] * ub[0] + ua[1] * ub[1]) + ua[2] * ub[2];
  if (s == 0.0) {
    if (c > 0.0) {
      for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    } else {
      double e[3] = {1.0, 0.0, 0.0};
      if (fabs(ua[0]) > fabs(ua[1]) || fabs(ua[0]) > fabs(ua[2])) {
        if (fabs(ua[1]) <= fabs(ua[2])) { e[0] = 0.0; e[1] = 1.0; }
        else { e[0] = 0.0; e[2] = 1.0; }
      }
      double n[3] = {ua[1] * e[2] - ua[2] * e[1], ua[2] * e[0] - ua[0] * e[2], ua[0] * e[1] - ua[1] * e[0]};
      double nn = norm3(n[0], n[1], n[2]);
      for (int i = 0; i < 3; ++i) n[i] /= nn;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = 2.0 * n[i] * n[j] - (i == j ? 1.0 : 0.0);
    }
  } else {
    double k[3] = {v[0] / s, v[1] / s, v[2] / s};
    double K[9] = {0.0, -k[2], k[1], k[2], 0.0, -k[0], -k[1], k[0], 0.0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        R[i * 3 + j] = (i == j ? c : 0.0) + s * K[i * 3 + j] + (1.0 - c) * k[i] * k[j];
  }
  for (int i = 0; i < 9; ++i) R[i] = cwt(R[i], abs_tol);
  return true;
}

// ------------------------------------------------------------------ truss
// Truss::create rotation (truss.rs:66-93)
__host__ __device__ inline bool truss_rotation(const double d[3], double abs_tol, double R[9]) {
  double L = norm3(d[0], d[1], d[2]);
  double t[3] = {L, 0.0, 0.0};
  return rotation_to_align(d, t, abs_tol, R);
}

// bar Jacobian at r = 0 (bar_2n_element_functions.rs:35-83): dx/dr = L/2
__host__ __device__ inline double bar_jacobian(double L) {
  double x1 = -1.0 * L / 2.0, x2 = L / 2.0;
  return ((0.0 - x1 * 0.5) + 0.0) + x2 * 0.5;
}

// ------------------------------------------------------------------ beam
// principal moments of inertia (beam.rs:119-160)
__host__ __device__ inline void beam_principal(double i11, double i22, double i12, double rel_tol, double& i11p,
                                               double& i22p, double& angle) {
  if (i11 != i22) {
    angle = atan(2.0 * i12 / (i22 - i11)) / 2.0;
  } else {
    double i11m, i22m;
    if (i22 < i11) { i11m = i11; i22m = (fabs(i22) - fabs(i22) * rel_tol) * i22 / fabs(i22); }
    else { i11m = (fabs(i11) - fabs(i11) * rel_tol) * i11 / fabs(i11); i22m = i22; }
    angle = atan(2.0 * i12 / (i22m - i11m)) / 2.0;
  }
  double c = cos(angle), s = sin(angle), s2 = sin(2.0 * angle);
  i11p = i11 * (c * c) + i22 * (s * s) - i12 * s2;
  i22p = i11 * (s * s) + i22 * (c * c) + i12 * s2;
  for (int i = 1; i11p < i22p && i <= 64; ++i) {
    angle = (atan(2.0 * i12 / (i22 - i11)) + PI_F32 * double(float(i))) / 2.0;
    c = cos(angle); s = sin(angle); s2 = sin(2.0 * angle);
    i11p = i11 * (c * c) + i22 * (s * s) - i12 * s2;
    i22p = i11 * (s * s) + i22 * (c * c) + i12 * s2;
  }
}

// beam direction cosines with the section roll (beam.rs:162-258)
__host__ __device__ inline bool beam_rotation(const double d[3], const double v1[3], double angle, double abs_tol,
                                              double out[9]) {
  double L = norm3(d[0], d[1], d[2]);
  double R0[9];
  double t[3] = {L, 0.0, 0.0};
  if (!rotation_to_align(d, t, abs_tol, R0)) return false;
  double f = ((v1[0] * d[0] + v1[1] * d[1]) + v1[2] * d[2]) / ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  double p[3] = {v1[0] - f * d[0], v1[1] - f * d[1], v1[2] - f * d[2]};
  double tp[3];
  for (int i = 0; i < 3; ++i) tp[i] = (R0[i * 3] * p[0] + R0[i * 3 + 1] * p[1]) + R0[i * 3 + 2] * p[2];
  double cosang = ((0.0 * tp[0] + 0.0 * tp[1]) + 1.0 * tp[2]) / (1.0 * norm3(tp[0], tp[1], tp[2]));
  double total = angle + acos(cosang);
  double cx = cwt(d[0] / L, abs_tol), cy = cwt(d[1] / L, abs_tol), cz = cwt(d[2] / L, abs_tol);
  double cxz = cwt(sqrt(cx * cx + cz * cz), abs_tol);
  double c = cwt(cos(total), abs_tol), s = cwt(sin(total), abs_tol);
  if (cxz != 0.0) {
    out[0] = cx; out[1] = cy; out[2] = cz;
    out[3] = (-1.0 * cx * cy * c - cz * s) / cxz;
    out[4] = cxz * c;
    out[5] = (-1.0 * cy * cz * c + cx * s) / cxz;
    out[6] = (cx * cy * s - cz * c) / cxz;
    out[7] = -1.0 * cxz * s;
    out[8] = (cy * cz * s + cx * c) / cxz;
  } else {
    out[0] = 0.0; out[1] = cy; out[2] = 0.0;
    out[3] = -1.0 * cy * c; out[4] = 0.0; out[5] = s;
    out[6] = cy * s; out[7] = 0.0; out[8] = c;
  }
  return true;
}

// ------------------------------------------------------------------ plate
// frame of a quadrilateral from nodes 2, 3, 4 (quadrilateral_4n_element_functions.rs:131-171)
__host__ __device__ inline bool quad_rotation(const double p2[3], const double p3[3], const double p4[3],
                                              double abs_tol, double R[9]) {
  double e34[3] = {p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
  double e32[3] = {p2[0] - p3[0], p2[1] - p3[1], p2[2] - p3[2]};
  double n[3] = {e34[1] * e32[2] - e34[2] * e32[1], e34[2] * e32[0] - e34[0] * e32[2],
                 e34[0] * e32[1] - e34[1] * e32[0]};
  double t[3] = {0.0, 0.0, norm3(n[0], n[1], n[2])};
  return rotation_to_align(n, t, abs_tol, R);
}

#pragma clang fp contract(fast)

// Per-integration-point kinematics of the 4-node plate in its local frame
// (node 3 at the origin): dh/dx, dh/dy (quad:613-653), det J (quad:477-503),
// MITC4 transverse-shear rows (plate.rs:392-511).
struct PlateIP {
  double w;        // det J * alpha
  double dx[4], dy[4];
  double sr[4][3];  // gamma_rz row: (w, theta_x, theta_y) per node
  double ss[4][3];  // gamma_sz row
};

// xl = {x1, y1, x2, y2, x4, y4}: local coordinates of nodes 1, 2, 4 (node 3 = origin)
__host__ __device__ inline void plate_ip(const double xl[6], double r, double s, PlateIP& o) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double q = 0.25;
  // dx/dr etc. (quad:252-338) with x3 = y3 = 0
  double x1q = x1 * q, x2q = x2 * q, x4q = x4 * q, y1q = y1 * q, y2q = y2 * q, y4q = y4 * q;
  double j00 = ((((x1q + x1q * s) - x2q) - x2q * s) + 0.0 * s) + x4q - x4q * s;
  double j01 = ((((y1q + y1q * s) - y2q) - y2q * s) + 0.0 * s) + y4q - y4q * s;
  double j10 = ((((x1q + x1q * r) + x2q) - x2q * r) - x4q) - x4q * r;
  double j11 = ((((y1q + y1q * r) + y2q) - y2q * r) - y4q) - y4q * r;
  double det = j00 * j11 - j01 * j10;
  double i00 = j11 / det, i01 = -1.0 * j01 / det, i10 = -1.0 * j10 / det, i11 = j00 / det;
  // shape-function derivatives (quad:505-583)
  double dr[4] = {q + q * s, -(q + q * s), -(q - q * s), q - q * s};
  double ds[4] = {q + q * r, q - q * r, -(q - q * r), -(q + q * r)};
  for (int a = 0; a < 4; ++a) {
    o.dx[a] = i00 * dr[a] + i01 * ds[a];
    o.dy[a] = i10 * dr[a] + i11 * ds[a];
  }
  o.w = det;  // alpha = 1 * 1 for the 2x2 rule
  double ax = x1 - x2 + x4, bx = x1 - x2 - x4, cx = x1 + x2 - x4;
  double ay = y1 - y2 + y4, by = y1 - y2 - y4, cy = y1 + y2 - y4;
  double mr = sqrt((cx + r * bx) * (cx + r * bx) + (cy + r * by) * (cy + r * by)) / (8.0 * det);
  double ms = sqrt((ax + s * bx) * (ax + s * bx) + (ay + s * by) * (ay + s * by)) / (8.0 * det);
  double ps = (1.0 + s), msn = (1.0 - s), pr = (1.0 + r), mrn = (1.0 - r);
  double tx12 = -(y1 - y2) / 4.0, ty12 = (x1 - x2) / 4.0;
  double tx43 = -(y4) / 4.0, ty43 = (x4) / 4.0;
  double tx14 = -(y1 - y4) / 4.0, ty14 = (x1 - x4) / 4.0;
  double tx23 = -(y2) / 4.0, ty23 = (x2) / 4.0;
  o.sr[0][0] = ps / 2.0 * mr;  o.sr[0][1] = ps * tx12 * mr;  o.sr[0][2] = ps * ty12 * mr;
  o.sr[1][0] = -ps / 2.0 * mr; o.sr[1][1] = ps * tx12 * mr;  o.sr[1][2] = ps * ty12 * mr;
  o.sr[2][0] = -msn / 2.0 * mr; o.sr[2][1] = msn * tx43 * mr; o.sr[2][2] = msn * ty43 * mr;
  o.sr[3][0] = msn / 2.0 * mr; o.sr[3][1] = msn * tx43 * mr; o.sr[3][2] = msn * ty43 * mr;
  o.ss[0][0] = pr / 2.0 * ms;  o.ss[0][1] = pr * tx14 * ms;  o.ss[0][2] = pr * ty14 * ms;
  o.ss[1][0] = mrn / 2.0 * ms; o.ss[1][1] = mrn * tx23 * ms; o.ss[1][2] = mrn * ty23 * ms;
  o.ss[2][0] = -mrn / 2.0 * ms; o.ss[2][1] = mrn * tx23 * ms; o.ss[2][2] = mrn * ty23 * ms;
  o.ss[3][0] = -pr / 2.0 * ms; o.ss[3][1] = pr * tx14 * ms;  o.ss[3][2] = pr * ty14 * ms;
}

struct PlateMat {
  double cm00, cm01, cm22;  // membrane: E t/(1-nu^2) * [1, nu, (1-nu)/2]
  double cb00, cb01, cb22;  // bending:  E t^3/(12(1-nu^2)) * [...]
  double cs;                // shear:    E t k/(2(1+nu)) (alpha = 1)
};
// The fused plate kernels must give the same bits for the same element, and
// whether a multiply and an add contract into an FMA depends on the code
// around an inlined helper.  So the plate helpers below are compiled without
// contraction, and the FMAs they want are written out.
#pragma clang fp contract(off)
// plate.rs:532-636
__host__ __device__ inline PlateMat plate_material(double E, double nu, double t, double k) {
  PlateMat m;
  double cm = E * t / (1.0 - nu * nu);
  double cb = E * (t * t * t) / (12.0 * (1.0 - nu * nu));
  double half = (1.0 - nu) / 2.0;
  m.cm00 = 1.0 * cm; m.cm01 = nu * cm; m.cm22 = half * cm;
  m.cb00 = 1.0 * cb; m.cb01 = nu * cb; m.cb22 = half * cb;
  m.cs = E * t * k / (2.0 * (1.0 + nu));
  return m;
}
#pragma clang fp contract(fast)

// Local-frame 6x6 block (p, q) of the plate stiffness accumulated over one
// integration point: membrane (u, v), bending (theta_x, theta_y) and MITC4
// shear (w, theta_x, theta_y) terms of plate.rs:513-672.
__host__ __device__ inline double sel4(double a, double b, double c, double d, int p) {
  return p == 0 ? a : (p == 1 ? b : (p == 2 ? c : d));
}

// Node data are picked with selects, not by runtime index: p and q vary per
// thread, and an array indexed at run time would sit in scratch memory.
__host__ __device__ inline void plate_block_add_ip(const PlateIP& g, const PlateMat& m, int p, int q,
                                                   double kb[36]) {
  const double w = g.w;
  const double dxp = sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p), dyp = sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
  const double dxq = sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], q), dyq = sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], q);
  double srp[3], srq[3], ssp[3], ssq[3];
  for (int i = 0; i < 3; ++i) {
    srp[i] = sel4(g.sr[0][i], g.sr[1][i], g.sr[2][i], g.sr[3][i], p);
    srq[i] = sel4(g.sr[0][i], g.sr[1][i], g.sr[2][i], g.sr[3][i], q);
    ssp[i] = sel4(g.ss[0][i], g.ss[1][i], g.ss[2][i], g.ss[3][i], p);
    ssq[i] = sel4(g.ss[0][i], g.ss[1][i], g.ss[2][i], g.ss[3][i], q);
  }
  // membrane: B_mem columns u_a = (dx_a, 0, dy_a), v_a = (0, dy_a, dx_a)
  double muu = m.cm00 * dxp * dxq + m.cm22 * dyp * dyq;
  double muv = m.cm01 * dxp * dyq + m.cm22 * dyp * dxq;
  double mvu = m.cm01 * dyp * dxq + m.cm22 * dxp * dyq;
  double mvv = m.cm00 * dyp * dyq + m.cm22 * dxp * dxq;
  // bending: B_bend columns tx_a = (0, dy_a, dx_a), ty_a = (-dx_a, 0, -dy_a)
  double bxx = m.cb00 * dyp * dyq + m.cb22 * dxp * dxq;
  double bxy = -(m.cb01 * dyp * dxq + m.cb22 * dxp * dyq);
  double byx = -(m.cb01 * dxp * dyq + m.cb22 * dyp * dxq);
  double byy = m.cb00 * dxp * dxq + m.cb22 * dyp * dyq;
  kb[0 * 6 + 0] += muu * w; kb[0 * 6 + 1] += muv * w;
  kb[1 * 6 + 0] += mvu * w; kb[1 * 6 + 1] += mvv * w;
  // shear: (w, tx, ty) x (w, tx, ty)
  const double cs = m.cs;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double sh = cs * (srp[i] * srq[j] + ssp[i] * ssq[j]);
      double be = 0.0;
      if (i == 1 && j == 1) be = bxx;
      else if (i == 1 && j == 2) be = bxy;
      else if (i == 2 && j == 1) be = byx;
      else if (i == 2 && j == 2) be = byy;
      kb[(2 + i) * 6 + 2 + j] += be * w + sh * w;
    }
}

// Local-frame row block of node p (columns of all four nodes q) of a flat
// plate, accumulated over one integration point.  Layout: block q at 13 q,
// entries (u,u) (u,v) | (v,u) (v,v) | (w,w) (w,tx) (w,ty) | (tx,w) (tx,tx)
// (tx,ty) | (ty,w) (ty,tx) (ty,ty).  The drilling entry (tz,tz) = KROT6 is
// handled by the caller.  Same per-entry arithmetic as plate_block_add_ip.
__host__ __device__ inline void plate_row_add_ip(const PlateIP& g, const PlateMat& m, int p, double acc[52]) {
  const double w = g.w;
  const double dxp = sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p);
  const double dyp = sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
  double srp[3], ssp[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    srp[i] = sel4(g.sr[0][i], g.sr[1][i], g.sr[2][i], g.sr[3][i], p);
    ssp[i] = sel4(g.ss[0][i], g.ss[1][i], g.ss[2][i], g.ss[3][i], p);
  }
  const double cs = m.cs;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double* o = acc + 13 * q;
    const double dxq = g.dx[q], dyq = g.dy[q];
    double muu = m.cm00 * dxp * dxq + m.cm22 * dyp * dyq;
    double muv = m.cm01 * dxp * dyq + m.cm22 * dyp * dxq;
    double mvu = m.cm01 * dyp * dxq + m.cm22 * dxp * dyq;
    double mvv = m.cm00 * dyp * dyq + m.cm22 * dxp * dxq;
    double bxx = m.cb00 * dyp * dyq + m.cb22 * dxp * dxq;
    double bxy = -(m.cb01 * dyp * dxq + m.cb22 * dxp * dyq);
    double byx = -(m.cb01 * dxp * dyq + m.cb22 * dyp * dxq);
    double byy = m.cb00 * dxp * dxq + m.cb22 * dyp * dyq;
    o[0] += muu * w; o[1] += muv * w; o[2] += mvu * w; o[3] += mvv * w;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sh = cs * (srp[i] * g.sr[q][j] + ssp[i] * g.ss[q][j]);
        double be = 0.0;
        if (i == 1 && j == 1) be = bxx;
        else if (i == 1 && j == 2) be = bxy;
        else if (i == 2 && j == 1) be = byx;
        else if (i == 2 && j == 2) be = byy;
        o[4 + 3 * i + j] += be * w + sh * w;
      }
  }
}

// Row block of node p of a flat plate (local frame), folded over the 2x2
// Gauss rule.  Every term of plate.rs:513-672 is a bilinear form in
// per-point shape-function data, so the four integration points reduce to six
// scalar integrals per column node q:
//   Mxx = sum w dx_p dx_q, Mxy = sum w dx_p dy_q, Myx = sum w dy_p dx_q, Myy = sum w dy_p dy_q
// and, because the MITC4 shear rows factor as B_r[a] = alpha_a(r,s) * v_a and
// B_s[a] = beta_a(r,s) * u_a with point-independent v_a, u_a (plate.rs:452-508),
//   A = sum w alpha_p alpha_q,  B = sum w beta_p beta_q.
// The 13 entries of block q follow from these with the constitutive
// constants.  Same values as the per-point sums up to rounding.
// xl = {x1, y1, x2, y2, x4, y4}: local coordinates relative to node 3.

#pragma clang fp contract(off)
// per-point data: w = det J (alpha = 1), dh/dx, dh/dy (quad:613-653), alpha, beta
struct PlatePoint {
  double w, dx[4], dy[4], al[4], be[4];
};
constexpr int PLATE_POINT_DOUBLES = 17;

// integration point ip in the crate's order (g,g), (-g,g), (-g,-g), (g,-g)
__host__ __device__ inline void plate_point(const double xl[6], int ip, PlatePoint& o) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double g = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? g : -g;
  const double s = (ip < 2) ? g : -g;
  const double q4 = 0.25;
  const double x1q = x1 * q4, x2q = x2 * q4, x4q = x4 * q4, y1q = y1 * q4, y2q = y2 * q4, y4q = y4 * q4;
  // Jacobian (quad:252-338 with x3 = y3 = 0), determinant and inverse (quad:380-503)
  const double j00 = ((((x1q + x1q * s) - x2q) - x2q * s) + x4q) - x4q * s;
  const double j01 = ((((y1q + y1q * s) - y2q) - y2q * s) + y4q) - y4q * s;
  const double j10 = ((((x1q + x1q * r) + x2q) - x2q * r) - x4q) - x4q * r;
  const double j11 = ((((y1q + y1q * r) + y2q) - y2q * r) - y4q) - y4q * r;
  const double det = j00 * j11 - j01 * j10;
  const double rdet = 1.0 / det;
  const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
  // shape-function derivatives (quad:505-583)
  const double dr[4] = {q4 + q4 * s, -(q4 + q4 * s), -(q4 - q4 * s), q4 - q4 * s};
  const double ds[4] = {q4 + q4 * r, q4 - q4 * r, -(q4 - q4 * r), -(q4 + q4 * r)};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    o.dx[a] = i00 * dr[a] + i01 * ds[a];
    o.dy[a] = i10 * dr[a] + i11 * ds[a];
  }
  // MITC4 multipliers gamma_rz, gamma_sz (plate.rs:424-447)
  const double ax = x1 - x2 + x4, bx = x1 - x2 - x4, cx = x1 + x2 - x4;
  const double ay = y1 - y2 + y4, by = y1 - y2 - y4, cy = y1 + y2 - y4;
  const double r8 = 0.125 * rdet;
  const double mr = sqrt((cx + r * bx) * (cx + r * bx) + (cy + r * by) * (cy + r * by)) * r8;
  const double ms = sqrt((ax + s * bx) * (ax + s * bx) + (ay + s * by) * (ay + s * by)) * r8;
  o.al[0] = (1.0 + s) * mr; o.al[1] = (1.0 + s) * mr; o.al[2] = (1.0 - s) * mr; o.al[3] = (1.0 - s) * mr;
  o.be[0] = (1.0 + r) * ms; o.be[1] = (1.0 - r) * ms; o.be[2] = (1.0 - r) * ms; o.be[3] = (1.0 + r) * ms;
  o.w = det;
}

// point-independent shear vectors (w, theta_x, theta_y components) of the
// gamma_rz rows (v) and gamma_sz rows (u), plate.rs:452-508
struct PlateShear {
  double v[4][3], u[4][3];
};
__host__ __device__ inline void plate_shear_vectors(const double xl[6], PlateShear& o) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double e12x = -(y1 - y2) * 0.25, e12y = (x1 - x2) * 0.25;
  const double e43x = -(y4) * 0.25, e43y = (x4) * 0.25;
  const double e14x = -(y1 - y4) * 0.25, e14y = (x1 - x4) * 0.25;
  const double e23x = -(y2) * 0.25, e23y = (x2) * 0.25;
  const double v[4][3] = {{0.5, e12x, e12y}, {-0.5, e12x, e12y}, {-0.5, e43x, e43y}, {0.5, e43x, e43y}};
  const double u[4][3] = {{0.5, e14x, e14y}, {0.5, e23x, e23y}, {-0.5, e23x, e23y}, {-0.5, e14x, e14y}};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) { o.v[a][i] = v[a][i]; o.u[a][i] = u[a][i]; }
}

// acc[24] = Mxx[4] Mxy[4] Myx[4] Myy[4] A[4] B[4] of row p; wdx = w dx_p, wdy = w dy_p,
// wal = w alpha_p, wbe = w beta_p of the same point
__host__ __device__ inline void plate_row_accumulate(const PlatePoint& g, double wdx, double wdy, double wal, double wbe,
                                                     double acc[24]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    acc[q] = __builtin_fma(wdx, g.dx[q], acc[q]);
    acc[4 + q] = __builtin_fma(wdx, g.dy[q], acc[4 + q]);
    acc[8 + q] = __builtin_fma(wdy, g.dx[q], acc[8 + q]);
    acc[12 + q] = __builtin_fma(wdy, g.dy[q], acc[12 + q]);
    acc[16 + q] = __builtin_fma(wal, g.al[q], acc[16 + q]);
    acc[20 + q] = __builtin_fma(wbe, g.be[q], acc[20 + q]);
  }
}

// the 13 entries of block q of row p from its six integrals; vp, up: shear
// vectors of node p scaled by the shear constant cs; vq, uq: those of node q
__host__ __device__ inline void plate_block_entries(double Mxx, double Mxy, double Myx, double Myy, double A, double B,
                                                    const PlateMat& m, const double vp[3], const double up[3],
                                                    const double vq[3], const double uq[3], double o[13]) {
  o[0] = __builtin_fma(m.cm00, Mxx, m.cm22 * Myy);
  o[1] = __builtin_fma(m.cm01, Mxy, m.cm22 * Myx);
  o[2] = __builtin_fma(m.cm01, Myx, m.cm22 * Mxy);
  o[3] = __builtin_fma(m.cm00, Myy, m.cm22 * Mxx);
  const double bxx = __builtin_fma(m.cb00, Myy, m.cb22 * Mxx);
  const double bxy = -__builtin_fma(m.cb01, Myx, m.cb22 * Mxy);
  const double byx = -__builtin_fma(m.cb01, Mxy, m.cb22 * Myx);
  const double byy = __builtin_fma(m.cb00, Mxx, m.cb22 * Myy);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sh = __builtin_fma(A * vp[i], vq[j], (B * up[i]) * uq[j]);
      double b = 0.0;
      if (i == 1 && j == 1) b = bxx;
      else if (i == 1 && j == 2) b = bxy;
      else if (i == 2 && j == 1) b = byx;
      else if (i == 2 && j == 2) b = byy;
      o[4 + 3 * i + j] = b + sh;
    }
}

// the 13 entries of each block q of row p (block q starts at out[BS * q]);
// vp, up: shear vectors of node p scaled by the shear constant cs
template <int BS>
__host__ __device__ inline void plate_row_finish(const double acc[24], const PlateMat& m, const double vp[3],
                                                 const double up[3], const PlateShear& sv, double* out) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    plate_block_entries(acc[q], acc[4 + q], acc[8 + q], acc[12 + q], acc[16 + q], acc[20 + q], m, vp, up, sv.v[q],
                        sv.u[q], out + BS * q);
  }
}

// single-lane version: row block of node p (used by the generic path)
template <int BS = 13>
__host__ __device__ inline void plate_row_folded(const double xl[6], const PlateMat& m, int p, double* out) {
  double acc[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) acc[k] = 0.0;
#pragma unroll 1
  for (int ip = 0; ip < 4; ++ip) {
    PlatePoint g;
    plate_point(xl, ip, g);
    const double wdx = g.w * sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p);
    const double wdy = g.w * sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
    const double wal = g.w * sel4(g.al[0], g.al[1], g.al[2], g.al[3], p);
    const double wbe = g.w * sel4(g.be[0], g.be[1], g.be[2], g.be[3], p);
    plate_row_accumulate(g, wdx, wdy, wal, wbe, acc);
  }
  PlateShear sv;
  plate_shear_vectors(xl, sv);
  double vp[3], up[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vp[i] = m.cs * sel4(sv.v[0][i], sv.v[1][i], sv.v[2][i], sv.v[3][i], p);
    up[i] = m.cs * sel4(sv.u[0][i], sv.u[1][i], sv.u[2][i], sv.u[3][i], p);
  }
  plate_row_finish<BS>(acc, m, vp, up, sv, out);
}
#pragma clang fp contract(fast)

// Gauss abscissa of the 2x2 rule in the crate's point order (g,g), (-g,g), (-g,-g), (g,-g)
__host__ __device__ inline double plate_gauss() { return sqrt(double(ONE_THIRD_F32)); }

// G = R6^T kb R6 with R6 = diag(R, R), in place (no-op when R is the identity)
__host__ __device__ inline void rotate_block(double kb[36], const double R[9]) {
  bool ident = R[0] == 1.0 && R[4] == 1.0 && R[8] == 1.0 && R[1] == 0.0 && R[2] == 0.0 && R[3] == 0.0 &&
               R[5] == 0.0 && R[6] == 0.0 && R[7] == 0.0;
  if (ident) return;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj) {
      double A[9], T[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i * 3 + j] = kb[(3 * bi + i) * 6 + 3 * bj + j];
      // T = R^T A
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i * 3 + j] = (R[0 * 3 + i] * A[0 * 3 + j] + R[1 * 3 + i] * A[1 * 3 + j]) + R[2 * 3 + i] * A[2 * 3 + j];
      // kb = T R
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          kb[(3 * bi + i) * 6 + 3 * bj + j] = (T[i * 3 + 0] * R[0 * 3 + j] + T[i * 3 + 1] * R[1 * 3 + j]) + T[i * 3 + 2] * R[2 * 3 + j];
    }
}

}  // namespace fem
