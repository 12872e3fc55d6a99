// plate_adjoint.hpp -- reverse-mode derivative of a plate's contribution to a
// scalar L(K) (the adjoint of plate.rs:160-724 as folded in element_math.hpp).
//
// Given the local-frame gradient Gl = T (dL/dK_e) T^T of one element, lane p
// (row block of node p) produces its additive share of
//   * P[7]   dL/d(cm00, cm01, cm22, cb00, cb01, cb22, cs): the constitutive
//            constants, from which E, nu, t and k derivatives follow;
//   * gxl[6] dL/d(x1, y1, x2, y2, x4, y4): local nodal coordinates at fixed frame;
//   * hx, hy dL/d(omega_x, omega_y): an infinitesimal rotation of the element
//            frame, dR = [omega]x R.  K_e does not depend on the in-plane gauge
//            of the frame, so omega_z never contributes.
// plate_adjoint_finish turns the sums of the four lanes into parameter and
// world-frame nodal-coordinate gradients.  No K_e is stored: the forward
// integrals are recomputed.
#pragma once
#include <hip/hip_runtime.h>

#include "element_math.hpp"

namespace fem {

// Reverse sweep of one integration point: the derivative with respect to xl of
//   l = sum_q w [dx_p (cxx_q dx_q + cxy_q dy_q) + dy_p (cyx_q dx_q + cyy_q dy_q)
//               + al_p cA_q al_q + be_p cB_q be_q]
// with c = {cxx[4], cxy[4], cyx[4], cyy[4], cA[4], cB[4]} held fixed, where
// w, dx, dy, al, be are the per-point data of plate_point (w = det J).  It
// runs in two parts.  The second is linear in the seven adjoints of the
// point's forward quantities (det, the inverse Jacobian, the two MITC4
// multipliers), so the four row lanes of an element add their seven values
// first and run the second part once per point.
// Part 1: the seven adjoints s7 = (g_det, g_i00, g_i01, g_i10, g_i11, g_mr,
// g_ms) of row p at point ip, from that point's data g (plate_point).
__host__ __device__ inline void plate_point_scalars(const PlatePoint& g, int ip, int p, const double c[24],
                                                    double s7[7]) {
  const double gs = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? gs : -gs;
  const double s = (ip < 2) ? gs : -gs;
  const double q4 = 0.25;
  const double pr = 1.0 + r, mr_ = 1.0 - r, ps = 1.0 + s, ms_ = 1.0 - s;
  const double dr[4] = {q4 + q4 * s, -(q4 + q4 * s), -(q4 - q4 * s), q4 - q4 * s};
  const double ds[4] = {q4 + q4 * r, q4 - q4 * r, -(q4 - q4 * r), -(q4 + q4 * r)};
  const double fa[4] = {ps, ps, ms_, ms_}, fb[4] = {pr, mr_, mr_, pr};
  const double det = g.w;
  const double dxp = sel4(g.dx[0], g.dx[1], g.dx[2], g.dx[3], p), dyp = sel4(g.dy[0], g.dy[1], g.dy[2], g.dy[3], p);
  const double alp = sel4(g.al[0], g.al[1], g.al[2], g.al[3], p), bep = sel4(g.be[0], g.be[1], g.be[2], g.be[3], p);
  double SX = 0.0, SY = 0.0, SA = 0.0, SB = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    SX += c[q] * g.dx[q] + c[4 + q] * g.dy[q];
    SY += c[8 + q] * g.dx[q] + c[12 + q] * g.dy[q];
    SA += c[16 + q] * g.al[q];
    SB += c[20 + q] * g.be[q];
  }
  double g_det = ((dxp * SX + dyp * SY) + alp * SA) + bep * SB;
  double g_i00 = 0.0, g_i01 = 0.0, g_i10 = 0.0, g_i11 = 0.0, g_mr = 0.0, g_ms = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const bool self = a == p;
    const double gdx = det * ((dxp * c[a] + dyp * c[8 + a]) + (self ? SX : 0.0));
    const double gdy = det * ((dxp * c[4 + a] + dyp * c[12 + a]) + (self ? SY : 0.0));
    const double gal = det * (alp * c[16 + a] + (self ? SA : 0.0));
    const double gbe = det * (bep * c[20 + a] + (self ? SB : 0.0));
    g_i00 += gdx * dr[a];
    g_i01 += gdx * ds[a];
    g_i10 += gdy * dr[a];
    g_i11 += gdy * ds[a];
    g_mr += gal * fa[a];
    g_ms += gbe * fb[a];
  }
  s7[0] = g_det; s7[1] = g_i00; s7[2] = g_i01; s7[3] = g_i10; s7[4] = g_i11; s7[5] = g_mr; s7[6] = g_ms;
}

// Part 2: from the seven adjoints of point ip (summed over rows) to xl.
__host__ __device__ inline void plate_point_reverse(const double xl[6], int ip, const double s7[7], double gxl[6]) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double gs = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? gs : -gs;
  const double s = (ip < 2) ? gs : -gs;
  const double q4 = 0.25;
  const double pr = 1.0 + r, mr_ = 1.0 - r, ps = 1.0 + s, ms_ = 1.0 - s;
  const double usx = (x1 * ps - x2 * ps) + x4 * ms_, usy = (y1 * ps - y2 * ps) + y4 * ms_;
  const double urx = (x1 * pr + x2 * mr_) - x4 * pr, ury = (y1 * pr + y2 * mr_) - y4 * pr;
  const double j00 = usx * q4, j01 = usy * q4, j10 = urx * q4, j11 = ury * q4;
  const double det = j00 * j11 - j01 * j10;
  const double rdet = 1.0 / det;
  const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
  const double sqr = sqrt(urx * urx + ury * ury), sqs = sqrt(usx * usx + usy * usy);
  const double m_r = sqr * 0.125 * rdet, m_s = sqs * 0.125 * rdet;
  double g_det = s7[0];
  const double g_i00 = s7[1], g_i01 = s7[2], g_i10 = s7[3], g_i11 = s7[4], g_mr = s7[5], g_ms = s7[6];
  g_det -= (g_mr * m_r + g_ms * m_s) * rdet;
  const double g_sqr = g_mr * 0.125 * rdet, g_sqs = g_ms * 0.125 * rdet;
  double g_j11 = g_i00 * rdet, g_j01 = -g_i01 * rdet, g_j10 = -g_i10 * rdet, g_j00 = g_i11 * rdet;
  g_det -= (((g_i00 * i00 + g_i01 * i01) + g_i10 * i10) + g_i11 * i11) * rdet;
  g_j00 += g_det * j11;
  g_j11 += g_det * j00;
  g_j01 -= g_det * j10;
  g_j10 -= g_det * j01;
  const double kux = g_j00 * q4 + (sqs > 0.0 ? g_sqs * usx / sqs : 0.0);
  const double kuy = g_j01 * q4 + (sqs > 0.0 ? g_sqs * usy / sqs : 0.0);
  const double krx = g_j10 * q4 + (sqr > 0.0 ? g_sqr * urx / sqr : 0.0);
  const double kry = g_j11 * q4 + (sqr > 0.0 ? g_sqr * ury / sqr : 0.0);
  gxl[0] += kux * ps + krx * pr;
  gxl[1] += kuy * ps + kry * pr;
  gxl[2] += -kux * ps + krx * mr_;
  gxl[3] += -kuy * ps + kry * mr_;
  gxl[4] += kux * ms_ - krx * pr;
  gxl[5] += kuy * ms_ - kry * pr;
}

// plate_point (element_math.hpp) that also returns the reciprocals the
// reverse sweep of the point needs (plate_point_reverse_r)
__host__ __device__ inline void plate_point_x(const double xl[6], int ip, PlatePoint& o, double& rdet_o,
                                              double& rqr, double& rqs) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double g = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? g : -g;
  const double s = (ip < 2) ? g : -g;
  const double q4 = 0.25;
  const double pr = 1.0 + r, mr_ = 1.0 - r, ps = 1.0 + s, ms_ = 1.0 - s;
  const double usx = (x1 * ps - x2 * ps) + x4 * ms_, usy = (y1 * ps - y2 * ps) + y4 * ms_;
  const double urx = (x1 * pr + x2 * mr_) - x4 * pr, ury = (y1 * pr + y2 * mr_) - y4 * pr;
  const double j00 = usx * q4, j01 = usy * q4, j10 = urx * q4, j11 = ury * q4;
  const double det = j00 * j11 - j01 * j10;
  const double rdet = 1.0 / det;
  const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
  const double dr[4] = {q4 * ps, -q4 * ps, -q4 * ms_, q4 * ms_};
  const double ds[4] = {q4 * pr, q4 * mr_, -q4 * mr_, -q4 * pr};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    o.dx[a] = i00 * dr[a] + i01 * ds[a];
    o.dy[a] = i10 * dr[a] + i11 * ds[a];
  }
  const double sqr = sqrt(urx * urx + ury * ury), sqs = sqrt(usx * usx + usy * usy);
  const double mr = sqr * 0.125 * rdet, ms = sqs * 0.125 * rdet;
  o.al[0] = ps * mr; o.al[1] = o.al[0]; o.al[2] = ms_ * mr; o.al[3] = o.al[2];
  o.be[0] = pr * ms; o.be[1] = mr_ * ms; o.be[2] = o.be[1]; o.be[3] = o.be[0];
  o.w = det;
  rdet_o = rdet;
  rqr = sqr > 0.0 ? 1.0 / sqr : 0.0;
  rqs = sqs > 0.0 ? 1.0 / sqs : 0.0;
}

// plate_point_reverse with the point's reciprocals given: rdet = 1 / det J and
// rqr, rqs = 1 / |(cx + r bx, cy + r by)|, 1 / |(ax + s bx, ay + s by)| (the
// MITC4 lengths; 0 where the length is 0), as the forward point formed them
__host__ __device__ inline void plate_point_reverse_r(const double xl[6], int ip, const double s7[7], double rdet,
                                                      double rqr, double rqs, double gxl[6]) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double gs = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? gs : -gs;
  const double s = (ip < 2) ? gs : -gs;
  const double q4 = 0.25;
  const double pr = 1.0 + r, mr_ = 1.0 - r, ps = 1.0 + s, ms_ = 1.0 - s;
  const double usx = (x1 * ps - x2 * ps) + x4 * ms_, usy = (y1 * ps - y2 * ps) + y4 * ms_;
  const double urx = (x1 * pr + x2 * mr_) - x4 * pr, ury = (y1 * pr + y2 * mr_) - y4 * pr;
  const double j00 = usx * q4, j01 = usy * q4, j10 = urx * q4, j11 = ury * q4;
  const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
  const double sqr = (urx * urx + ury * ury) * rqr, sqs = (usx * usx + usy * usy) * rqs;  // the lengths
  const double m_r = sqr * 0.125 * rdet, m_s = sqs * 0.125 * rdet;
  double g_det = s7[0];
  const double g_i00 = s7[1], g_i01 = s7[2], g_i10 = s7[3], g_i11 = s7[4], g_mr = s7[5], g_ms = s7[6];
  g_det -= (g_mr * m_r + g_ms * m_s) * rdet;
  const double g_sqr = g_mr * 0.125 * rdet, g_sqs = g_ms * 0.125 * rdet;
  double g_j11 = g_i00 * rdet, g_j01 = -g_i01 * rdet, g_j10 = -g_i10 * rdet, g_j00 = g_i11 * rdet;
  g_det -= (((g_i00 * i00 + g_i01 * i01) + g_i10 * i10) + g_i11 * i11) * rdet;
  g_j00 += g_det * j11;
  g_j11 += g_det * j00;
  g_j01 -= g_det * j10;
  g_j10 -= g_det * j01;
  const double kux = g_j00 * q4 + g_sqs * usx * rqs;
  const double kuy = g_j01 * q4 + g_sqs * usy * rqs;
  const double krx = g_j10 * q4 + g_sqr * urx * rqr;
  const double kry = g_j11 * q4 + g_sqr * ury * rqr;
  gxl[0] += kux * ps + krx * pr;
  gxl[1] += kuy * ps + kry * pr;
  gxl[2] += -kux * ps + krx * mr_;
  gxl[3] += -kuy * ps + kry * mr_;
  gxl[4] += kux * ms_ - krx * pr;
  gxl[5] += kuy * ms_ - kry * pr;
}

// adjoint of the shear vectors of node a (components 1 and 2; component 0 is
// the constant +-1/2): gv = dL/dv_a, gu = dL/du_a, mapped onto xl through
//   e12 = (-(y1-y2), x1-x2)/4, e43 = (-y4, x4)/4, e14 = (-(y1-y4), x1-x4)/4, e23 = (-y2, x2)/4
__host__ __device__ inline void plate_shear_vector_adjoint(int a, double gv1, double gv2, double gu1, double gu2,
                                                           double gxl[6]) {
  const double q = 0.25;
  if (a <= 1) {  // v_a carries e12
    gxl[1] -= gv1 * q; gxl[3] += gv1 * q;
    gxl[0] += gv2 * q; gxl[2] -= gv2 * q;
  } else {       // e43
    gxl[5] -= gv1 * q;
    gxl[4] += gv2 * q;
  }
  if (a == 0 || a == 3) {  // u_a carries e14
    gxl[1] -= gu1 * q; gxl[5] += gu1 * q;
    gxl[0] += gu2 * q; gxl[4] -= gu2 * q;
  } else {                 // e23
    gxl[3] -= gu1 * q;
    gxl[2] += gu2 * q;
  }
}

// h += antisymmetric parts (omega_x, omega_y) of K G^T + K^T G for one pair of
// 3x3 sub-blocks K = k[r0.., c0..], G = gl[r0.., c0..] of 6x6 row-major blocks
__host__ __device__ inline void plate_frame_adjoint_sub(const double* k, const double* gl, int r0, int c0, double& hx,
                                                        double& hy) {
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const double k0c = k[(r0 + 0) * 6 + c0 + cc], k1c = k[(r0 + 1) * 6 + c0 + cc], k2c = k[(r0 + 2) * 6 + c0 + cc];
    const double g0c = gl[(r0 + 0) * 6 + c0 + cc], g1c = gl[(r0 + 1) * 6 + c0 + cc], g2c = gl[(r0 + 2) * 6 + c0 + cc];
    const double kc0 = k[(r0 + cc) * 6 + c0 + 0], kc1 = k[(r0 + cc) * 6 + c0 + 1], kc2 = k[(r0 + cc) * 6 + c0 + 2];
    const double gc0 = gl[(r0 + cc) * 6 + c0 + 0], gc1 = gl[(r0 + cc) * 6 + c0 + 1], gc2 = gl[(r0 + cc) * 6 + c0 + 2];
    hx += (k2c * g1c - k1c * g2c) + (kc2 * gc1 - kc1 * gc2);
    hy += (k0c * g2c - k2c * g0c) + (kc0 * gc2 - kc2 * gc0);
  }
}

// Per-lane share of the adjoint sums for block q of row p.  Mq: the six
// integrals of column q (Mxx, Mxy, Myx, Myy, A, B; plate_row_accumulate),
// vp/up and vq/uq: the unscaled shear vectors of nodes p and q, gl: the
// local-frame gradient block (p, q), 6x6 row-major.  cq receives the integral
// coefficients of column q (layout of plate_point_scalars: cxx, cxy, cyx, cyy,
// cA, cB).
__host__ __device__ inline void plate_block_adjoint_q(const double Mq[6], const PlateMat& m, const double vp[3],
                                                      const double up[3], const double vq[3], const double uq[3],
                                                      int p, int q, const double gl[36], double P[7], double cq[6],
                                                      double gxl[6], double& hx, double& hy, bool frame = true) {
  const double Mxx = Mq[0], Mxy = Mq[1], Myx = Mq[2], Myy = Mq[3], A = Mq[4], B = Mq[5];
  const double Guu = gl[0], Guv = gl[1], Gvu = gl[6], Gvv = gl[7];
  const double Gxx = gl[21], Gxy = gl[22], Gyx = gl[27], Gyy = gl[28];
  // constitutive partials
  P[0] += Guu * Mxx + Gvv * Myy;
  P[1] += Guv * Mxy + Gvu * Myx;
  P[2] += ((Guu * Myy + Guv * Myx) + Gvu * Mxy) + Gvv * Mxx;
  P[3] += Gxx * Myy + Gyy * Mxx;
  P[4] -= Gxy * Myx + Gyx * Mxy;
  P[5] += ((Gxx * Mxx - Gxy * Mxy) - Gyx * Myx) + Gyy * Myy;
  // shear: Gs = gl[2..4][2..4]; v_p^T Gs v_q, u_p^T Gs u_q and the one-sided products
  double Gv_q[3], Gu_q[3], vG[3], uG[3];  // Gs v_q, Gs u_q, v_p^T Gs, u_p^T Gs
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Gv_q[i] = 0.0; Gu_q[i] = 0.0; vG[i] = 0.0; uG[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double gs = gl[(2 + i) * 6 + 2 + j];
      Gv_q[i] += gs * vq[j];
      Gu_q[i] += gs * uq[j];
      vG[j] += vp[i] * gs;
      uG[j] += up[i] * gs;
    }
  const double sv_ = (vp[0] * Gv_q[0] + vp[1] * Gv_q[1]) + vp[2] * Gv_q[2];
  const double su_ = (up[0] * Gu_q[0] + up[1] * Gu_q[1]) + up[2] * Gu_q[2];
  P[6] += A * sv_ + B * su_;
  // integral coefficients of column q
  cq[0] = ((m.cm00 * Guu + m.cm22 * Gvv) + m.cb22 * Gxx) + m.cb00 * Gyy;
  cq[1] = ((m.cm01 * Guv + m.cm22 * Gvu) - m.cb22 * Gxy) - m.cb01 * Gyx;
  cq[2] = ((m.cm22 * Guv + m.cm01 * Gvu) - m.cb01 * Gxy) - m.cb22 * Gyx;
  cq[3] = ((m.cm22 * Guu + m.cm00 * Gvv) + m.cb00 * Gxx) + m.cb22 * Gyy;
  cq[4] = m.cs * sv_;
  cq[5] = m.cs * su_;
  // shear vectors: d/dv_p = cs A Gs v_q, d/dv_q = cs A v_p^T Gs (same with u, B)
  const double sa = m.cs * A, sb = m.cs * B;
  plate_shear_vector_adjoint(p, sa * Gv_q[1], sa * Gv_q[2], sb * Gu_q[1], sb * Gu_q[2], gxl);
  plate_shear_vector_adjoint(q, sa * vG[1], sa * vG[2], sb * uG[1], sb * uG[2], gxl);
  // frame rotation: needs the local block values (skipped by callers whose
  // gradient blocks carry no in-plane / out-of-plane coupling: then it is 0)
  if (!frame) return;
  double kb[36];
#pragma unroll
  for (int i = 0; i < 36; ++i) kb[i] = 0.0;
  double o[13];
  double vps[3], ups[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vps[i] = m.cs * vp[i];
    ups[i] = m.cs * up[i];
  }
  plate_block_entries(Mxx, Mxy, Myx, Myy, A, B, m, vps, ups, vq, uq, o);
  kb[0] = o[0]; kb[1] = o[1]; kb[6] = o[2]; kb[7] = o[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) kb[(2 + i) * 6 + 2 + j] = o[4 + 3 * i + j];
  if (p == q) kb[35] = KROT6;
  plate_frame_adjoint_sub(kb, gl, 0, 0, hx, hy);
  plate_frame_adjoint_sub(kb, gl, 0, 3, hx, hy);
  plate_frame_adjoint_sub(kb, gl, 3, 0, hx, hy);
  plate_frame_adjoint_sub(kb, gl, 3, 3, hx, hy);
}

// Same for block q of row p given the whole row's integrals acc[24] and the
// element's shear vectors; writes column q of c[24].
__host__ __device__ inline void plate_block_adjoint(const double acc[24], const PlateMat& m, const PlateShear& sv,
                                                    int p, int q, const double gl[36], double P[7], double c[24],
                                                    double gxl[6], double& hx, double& hy, bool frame = true) {
  const double Mq[6] = {acc[q], acc[4 + q], acc[8 + q], acc[12 + q], acc[16 + q], acc[20 + q]};
  double vp[3], up[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vp[i] = sel4(sv.v[0][i], sv.v[1][i], sv.v[2][i], sv.v[3][i], p);
    up[i] = sel4(sv.u[0][i], sv.u[1][i], sv.u[2][i], sv.u[3][i], p);
  }
  double cq[6];
  plate_block_adjoint_q(Mq, m, vp, up, sv.v[q], sv.u[q], p, q, gl, P, cq, gxl, hx, hy, frame);
#pragma unroll
  for (int k = 0; k < 6; ++k) c[4 * k + q] = cq[k];
}

// Element-level finish from the four lanes' sums.
//   loc[a][k]: local coordinates R (X_a - X_3) of the crate's nodes 1..4 (a = 0..3)
//   grad_param = dL/d(E, nu, t, k); gX[a][k] = dL/dX_a (world frame)
__host__ __device__ inline void plate_adjoint_finish(const double P[7], const double gxl_in[6], double hx, double hy,
                                                     const double loc[4][3], const double R[9], double E, double nu,
                                                     double t, double k, double grad_param[4], double gX[4][3]) {
  // parameters: cm = E t/(1-nu^2), cb = E t^3/(12 (1-nu^2)), cs = E t k/(2 (1+nu))
  const double cm = E * t / (1.0 - nu * nu);
  const double cb = E * (t * t * t) / (12.0 * (1.0 - nu * nu));
  const double cs = E * t * k / (2.0 * (1.0 + nu));
  const double h = (1.0 - nu) / 2.0;
  const double Lm = cm * ((P[0] + nu * P[1]) + h * P[2]);
  const double Lb = cb * ((P[3] + nu * P[4]) + h * P[5]);
  const double Ls = cs * P[6];
  const double dcm = cm * 2.0 * nu / (1.0 - nu * nu), dcb = cb * 2.0 * nu / (1.0 - nu * nu);
  grad_param[0] = ((Lm + Lb) + Ls) / E;
  grad_param[1] = (((P[0] * dcm + P[1] * (cm + nu * dcm)) + P[2] * (h * dcm - 0.5 * cm)) +
                   ((P[3] * dcb + P[4] * (cb + nu * dcb)) + P[5] * (h * dcb - 0.5 * cb))) -
                  P[6] * cs / (1.0 + nu);
  grad_param[2] = ((Lm + 3.0 * Lb) + Ls) / t;
  grad_param[3] = Ls / k;
  // frame rotation: local z of nodes 1, 2, 4 feed back through dxl = (omega x p)_{x,y}
  double gxl[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) gxl[i] = gxl_in[i];
  const int nidx[3] = {0, 1, 3};
  double gwx = hx, gwy = hy;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double z = loc[nidx[m]][2];
    gwx -= gxl[2 * m + 1] * z;
    gwy += gxl[2 * m] * z;
  }
  // omega from the normal n = (X4 - X3) x (X2 - X3): omega_x = n~_y/|n|, omega_y = -n~_x/|n|
  // with n~ = R dn = da41 x p2 + p4 x da21 (local frame)
  const double* p2 = loc[1];
  const double* p4 = loc[3];
  const double nx = p4[1] * p2[2] - p4[2] * p2[1], ny = p4[2] * p2[0] - p4[0] * p2[2], nz = p4[0] * p2[1] - p4[1] * p2[0];
  const double nn = sqrt((nx * nx + ny * ny) + nz * nz);
  const double gnx = -gwy / nn, gny = gwx / nn;
  double ga[4][3];
  ga[0][0] = gxl[0]; ga[0][1] = gxl[1]; ga[0][2] = 0.0;
  ga[1][0] = gxl[2]; ga[1][1] = gxl[3]; ga[1][2] = 0.0;
  ga[3][0] = gxl[4]; ga[3][1] = gxl[5]; ga[3][2] = 0.0;
  // d n~_x = da41_y z2 - da41_z y2 + y4 da21_z - z4 da21_y
  // d n~_y = da41_z x2 - da41_x z2 + z4 da21_x - x4 da21_z
  const double g41[3] = {-gny * p2[2], gnx * p2[2], -gnx * p2[1] + gny * p2[0]};
  const double g21[3] = {gny * p4[2], -gnx * p4[2], gnx * p4[1] - gny * p4[0]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ga[3][i] += g41[i];
    ga[1][i] += g21[i];
    ga[2][i] = -((ga[0][i] + ga[1][i]) + ga[3][i]);
  }
  // world frame: dL/dX = R^T dL/da
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) gX[a][i] = (R[0 * 3 + i] * ga[a][0] + R[1 * 3 + i] * ga[a][1]) + R[2 * 3 + i] * ga[a][2];
}

// plate_adjoint_finish for a plate whose frame is the identity and whose nodes
// lie in z = 0 with no frame-rotation sums (hx = hy = 0): the rotation term
// vanishes and the world gradient is the local one.  The constants' reciprocals
// come from two divisions.
__host__ __device__ inline void plate_adjoint_finish_flat(const double P[7], const double gxl[6], double E, double nu,
                                                          double t, double k, double grad_param[4], double gX[4][3]) {
  const double inv_n = 1.0 / ((1.0 - nu) * (1.0 + nu));  // 1 / (1 - nu^2)
  const double inv_etk = 1.0 / ((E * t) * k);
  const double cm = E * t * inv_n;
  const double cb = E * (t * t * t) * (1.0 / 12.0) * inv_n;
  const double inv_1pn = (1.0 - nu) * inv_n;  // 1 / (1 + nu)
  const double cs = E * t * k * 0.5 * inv_1pn;
  const double h = (1.0 - nu) / 2.0;
  const double Lm = cm * ((P[0] + nu * P[1]) + h * P[2]);
  const double Lb = cb * ((P[3] + nu * P[4]) + h * P[5]);
  const double Ls = cs * P[6];
  const double dcm = cm * 2.0 * nu * inv_n, dcb = cb * 2.0 * nu * inv_n;
  grad_param[0] = ((Lm + Lb) + Ls) * (t * k * inv_etk);
  grad_param[1] = (((P[0] * dcm + P[1] * (cm + nu * dcm)) + P[2] * (h * dcm - 0.5 * cm)) +
                   ((P[3] * dcb + P[4] * (cb + nu * dcb)) + P[5] * (h * dcb - 0.5 * cb))) -
                  P[6] * cs * inv_1pn;
  grad_param[2] = ((Lm + 3.0 * Lb) + Ls) * (E * k * inv_etk);
  grad_param[3] = Ls * (E * t * inv_etk);
  const double g3x = -((gxl[0] + gxl[2]) + gxl[4]), g3y = -((gxl[1] + gxl[3]) + gxl[5]);
  gX[0][0] = gxl[0]; gX[0][1] = gxl[1];
  gX[1][0] = gxl[2]; gX[1][1] = gxl[3];
  gX[2][0] = g3x;    gX[2][1] = g3y;
  gX[3][0] = gxl[4]; gX[3][1] = gxl[5];
#pragma unroll
  for (int a = 0; a < 4; ++a) gX[a][2] = 0.0;
}

// Compliance by strain energy: the share of integration point ip in
// L = -u_l^T k u_l, with U[a] = (u, v, w, theta_x, theta_y, theta_z) the
// local-frame displacements of node a, plus the drilling term of node ip.
// Adds to P, gxl, hx, hy what plate_block_adjoint_q adds over the sixteen
// blocks for G = -u_p u_q^T, without forming a block: each block term is a
// bilinear form in nodal values, so its contraction with u u^T is the point's
// quadratic form in the strains
//   membrane eps = (ux, vy, uy + vx), bending kappa = (ay, -bx, ax - by)
//   (a = theta_x, b = theta_y), shear gamma_r = mr rho, gamma_s = ms sigma
// (rho, sigma: the MITC4 tying sums of the shear vectors against U), and
// every derivative follows from the strains, the stress resultants and the
// natural-coordinate derivatives of U:
//   * P from the squared strains;
//   * the point adjoints of plate_point_reverse: det from the energy density,
//     the inverse Jacobian from resultant x natural derivative, mr and ms
//     from the shear strains;
//   * the shear vectors from the tying sums;
//   * hx, hy = -2 sum_a (x_a x f_a)_{x,y} over the translation and rotation
//     triples, f = k u_l the local nodal forces (their point share here).
__host__ __device__ inline void plate_compliance_point(const double xl[6], int ip, const PlateMat& m,
                                                       const PlateShear& sv, const double U[4][6], double P[7],
                                                       double gxl[6], double& hx, double& hy) {
  const double x1 = xl[0], y1 = xl[1], x2 = xl[2], y2 = xl[3], x4 = xl[4], y4 = xl[5];
  const double gs = sqrt(double(ONE_THIRD_F32));
  const double r = (ip == 0 || ip == 3) ? gs : -gs;
  const double s = (ip < 2) ? gs : -gs;
  const double q4 = 0.25;
  const double pr = 1.0 + r, mr_ = 1.0 - r, ps = 1.0 + s, ms_ = 1.0 - s;
  const double usx = (x1 * ps - x2 * ps) + x4 * ms_, usy = (y1 * ps - y2 * ps) + y4 * ms_;
  const double urx = (x1 * pr + x2 * mr_) - x4 * pr, ury = (y1 * pr + y2 * mr_) - y4 * pr;
  const double j00 = usx * q4, j01 = usy * q4, j10 = urx * q4, j11 = ury * q4;
  const double det = j00 * j11 - j01 * j10;
  const double rdet = 1.0 / det;
  const double i00 = j11 * rdet, i01 = -j01 * rdet, i10 = -j10 * rdet, i11 = j00 * rdet;
  const double mr = sqrt(urx * urx + ury * ury) * 0.125 * rdet, ms = sqrt(usx * usx + usy * usy) * 0.125 * rdet;
  const double dr[4] = {q4 * ps, -q4 * ps, -q4 * ms_, q4 * ms_};
  const double ds[4] = {q4 * pr, q4 * mr_, -q4 * mr_, -q4 * pr};
  const double fa[4] = {ps, ps, ms_, ms_}, fb[4] = {pr, mr_, mr_, pr};
  // natural derivatives of u, v, theta_x, theta_y
  const int comp[4] = {0, 1, 3, 4};
  double fr[4], fs[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    fr[f] = ((dr[0] * U[0][comp[f]] + dr[1] * U[1][comp[f]]) + dr[2] * U[2][comp[f]]) + dr[3] * U[3][comp[f]];
    fs[f] = ((ds[0] * U[0][comp[f]] + ds[1] * U[1][comp[f]]) + ds[2] * U[2][comp[f]]) + ds[3] * U[3][comp[f]];
  }
  const double ux = i00 * fr[0] + i01 * fs[0], uy = i10 * fr[0] + i11 * fs[0];
  const double vx = i00 * fr[1] + i01 * fs[1], vy = i10 * fr[1] + i11 * fs[1];
  const double ax = i00 * fr[2] + i01 * fs[2], ay = i10 * fr[2] + i11 * fs[2];
  const double bx = i00 * fr[3] + i01 * fs[3], by = i10 * fr[3] + i11 * fs[3];
  // stress resultants: derivatives of the energy density phi with respect to
  // ux (Nux), vy (Nvy), uy and vx (Nxy), ay (May), bx (Mbx), ax (Mxy; by: -Mxy)
  const double Nux = 2.0 * (m.cm00 * ux + m.cm01 * vy);
  const double Nvy = 2.0 * (m.cm00 * vy + m.cm01 * ux);
  const double Nxy = 2.0 * m.cm22 * (uy + vx);
  const double May = 2.0 * (m.cb00 * ay - m.cb01 * bx);
  const double Mbx = 2.0 * (m.cb00 * bx - m.cb01 * ay);
  const double Mxy = 2.0 * m.cb22 * (ax - by);
  const double phi = 0.5 * (((Nux * ux + Nvy * vy) + Nxy * (uy + vx)) + ((May * ay + Mbx * bx) + Mxy * (ax - by)));
  // transverse shear: tying sums of the shear vectors against (w, theta_x, theta_y)
  double sp[4], tp[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    sp[a] = (sv.v[a][0] * U[a][2] + sv.v[a][1] * U[a][3]) + sv.v[a][2] * U[a][4];
    tp[a] = (sv.u[a][0] * U[a][2] + sv.u[a][1] * U[a][3]) + sv.u[a][2] * U[a][4];
  }
  const double rho = ps * (sp[0] + sp[1]) + ms_ * (sp[2] + sp[3]);
  const double sig = pr * (tp[0] + tp[3]) + mr_ * (tp[1] + tp[2]);
  const double gr = mr * rho, gz = ms * sig;
  const double w = det;
  // constitutive partials
  P[0] -= w * (ux * ux + vy * vy);
  P[1] -= w * (2.0 * ux * vy);
  P[2] -= w * ((uy + vx) * (uy + vx));
  P[3] -= w * (ay * ay + bx * bx);
  P[4] += w * (2.0 * ay * bx);
  P[5] -= w * ((ax - by) * (ax - by));
  P[6] -= w * (gr * gr + gz * gz);
  // point adjoints -> local coordinates
  const double cs2 = 2.0 * m.cs;
  double s7[7];
  s7[0] = -(phi + m.cs * (gr * gr + gz * gz));
  s7[1] = -w * (((Nux * fr[0] + Nxy * fr[1]) + Mxy * fr[2]) + Mbx * fr[3]);
  s7[2] = -w * (((Nux * fs[0] + Nxy * fs[1]) + Mxy * fs[2]) + Mbx * fs[3]);
  s7[3] = -w * (((Nxy * fr[0] + Nvy * fr[1]) + May * fr[2]) - Mxy * fr[3]);
  s7[4] = -w * (((Nxy * fs[0] + Nvy * fs[1]) + May * fs[2]) - Mxy * fs[3]);
  s7[5] = -w * cs2 * gr * rho;
  s7[6] = -w * cs2 * gz * sig;
  plate_point_reverse(xl, ip, s7, gxl);
  // shear vectors, and this point's share of the nodal forces f = k u_l
  const double cr = -w * cs2 * gr * mr, cz = -w * cs2 * gz * ms;  // dL/d(tying sums) per unit multiplier
  const double hw = 0.5 * w;
  double ex = 0.0, ey = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double dsp = cr * fa[a], dtp = cz * fb[a];
    plate_shear_vector_adjoint(a, dsp * U[a][3], dsp * U[a][4], dtp * U[a][3], dtp * U[a][4], gxl);
    const double dxa = i00 * dr[a] + i01 * ds[a], dya = i10 * dr[a] + i11 * ds[a];
    const double cv = (w * m.cs) * gr * (fa[a] * mr), cu = (w * m.cs) * gz * (fb[a] * ms);
    const double fu = hw * (Nux * dxa + Nxy * dya);
    const double fv = hw * (Nxy * dxa + Nvy * dya);
    const double fw = cv * sv.v[a][0] + cu * sv.u[a][0];
    const double fta = hw * (Mxy * dxa + May * dya) + (cv * sv.v[a][1] + cu * sv.u[a][1]);
    const double ftb = hw * (Mbx * dxa - Mxy * dya) + (cv * sv.v[a][2] + cu * sv.u[a][2]);
    ex += (U[a][1] * fw - U[a][2] * fv) - U[a][5] * ftb;
    ey += (U[a][2] * fu - U[a][0] * fw) + U[a][5] * fta;
  }
  // drilling force of node ip (KROT6 on its theta_z)
  const double tzx = ip == 0 ? U[0][3] : (ip == 1 ? U[1][3] : (ip == 2 ? U[2][3] : U[3][3]));
  const double tzy = ip == 0 ? U[0][4] : (ip == 1 ? U[1][4] : (ip == 2 ? U[2][4] : U[3][4]));
  const double tzz = ip == 0 ? U[0][5] : (ip == 1 ? U[1][5] : (ip == 2 ? U[2][5] : U[3][5]));
  const double fz = KROT6 * tzz;
  ex += tzy * fz;
  ey -= tzx * fz;
  hx -= 2.0 * ex;
  hy -= 2.0 * ey;
}

}  // namespace fem
