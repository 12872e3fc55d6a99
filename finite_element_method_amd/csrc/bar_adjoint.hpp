// bar_adjoint.hpp -- derivatives of a truss's or beam's contribution to a
// scalar L(K) with respect to its parameters (and, for trusses, its node
// coordinates), given dL/dK_e.  The stiffness is recomputed, never stored.
#pragma once
#include <hip/hip_runtime.h>

#include "element_math.hpp"

namespace fem {

// Truss (truss.rs:95-189): K_e = (area E / L) sigma_pq r r^T on the
// translational blocks, r = d / L, sigma = +1 on the diagonal blocks, -1 off.
// H = sum_pq sigma_pq G_pq (translational 3x3 blocks of dL/dK_e).
// Outputs dL/d(E, A, A2) (A2 = NaN: no taper, its slot gets 0) and dL/dd with
// d = X2 - X1 (node 1 receives -gd, node 2 +gd).
__host__ __device__ inline void truss_adjoint(const double d[3], const double r0[3], double E, double A, double A2,
                                              const double H[9], double gparam[3], double gd[3]) {
  const double L = norm3(d[0], d[1], d[2]);
  const bool taper = A2 == A2;
  const double area = taper ? (A2 - A) / 2.0 * 0.0 + A - (A2 - A) / 2.0 * -1.0 : A;
  // L_e with the frame row the forward uses
  double rHr = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) rHr += r0[i] * H[i * 3 + j] * r0[j];
  const double k00 = area * E / L;
  const double Le = k00 * rHr;
  gparam[0] = Le / E;
  const double garea = Le / area;
  gparam[1] = taper ? 0.5 * garea : garea;
  gparam[2] = taper ? 0.5 * garea : 0.0;
  // f(d) = area E d^T H d / |d|^3
  double Hd[3], dH[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Hd[i] = (H[i * 3 + 0] * d[0] + H[i * 3 + 1] * d[1]) + H[i * 3 + 2] * d[2];
    dH[i] = (d[0] * H[0 * 3 + i] + d[1] * H[1 * 3 + i]) + d[2] * H[2 * 3 + i];
  }
  const double dHd = (d[0] * Hd[0] + d[1] * Hd[1]) + d[2] * Hd[2];
  const double L2 = L * L, L3 = L2 * L;
  const double c = area * E / L3;
#pragma unroll
  for (int i = 0; i < 3; ++i) gd[i] = c * ((Hd[i] + dH[i]) - 3.0 * dHd * d[i] / L2);
}

// Beam (beam.rs:260-605): K_e = T^T (sum_t scale_t v_t v_t^T) T with six
// rank-one terms t = u, v, w, theta_u, theta_v, theta_w.  beam_term_sums adds
// S_t += v_t,p^T Gl_pq v_t,q for one local-frame 6x6 block (bp, bq: the
// scaled shape-function derivatives of nodes p and q).
__host__ __device__ inline void beam_term_sums(const double gl[36], double bp, double bq, double S[6]) {
  S[0] += bp * bq * gl[0 * 6 + 0];
  S[1] += ((bp * bq * gl[1 * 6 + 1] - 0.5 * bp * gl[1 * 6 + 5]) - 0.5 * bq * gl[5 * 6 + 1]) + 0.25 * gl[5 * 6 + 5];
  S[2] += ((bp * bq * gl[2 * 6 + 2] - 0.5 * bp * gl[2 * 6 + 4]) - 0.5 * bq * gl[4 * 6 + 2]) + 0.25 * gl[4 * 6 + 4];
  S[3] += bp * bq * gl[3 * 6 + 3];
  S[4] += bp * bq * gl[4 * 6 + 4];
  S[5] += bp * bq * gl[5 * 6 + 5];
}

// dL/d(E, nu, A, I11, I22, I12, It, k, v1x, v1y, v1z) from the term sums.
//   qj = det J * alpha (the forward's integration factor)
//   i11p, i22p, angle: principal moments and angle (beam.rs:119-160)
// I11 and I22 enter through the principal moments; with I12 = 0 the principal
// angle is stationary in them, so their derivatives are exact there.  The
// derivatives that rotate the section (I12, the local-axis direction, and
// I11/I22 when I12 != 0 or I11 == I22) are not provided: NaN.
__host__ __device__ inline void beam_adjoint_params(const double S[6], double qj, double E, double nu, double A,
                                                    double i11, double i22, double i12, double it, double k,
                                                    double i11p, double i22p, double angle, double gparam[11]) {
  const double G = E / (2.0 * (1.0 + nu));
  const double S12 = S[1] + S[2];
  gparam[0] = qj * ((((A * S[0] + (G / E) * A * k * S12) + (G / E) * it * S[3]) + i22p * S[4]) + i11p * S[5]);
  gparam[1] = qj * (-G / (1.0 + nu)) * (A * k * S12 + it * S[3]);
  gparam[2] = qj * (E * S[0] + G * k * S12);
  const double nan = __builtin_nan("");
  if (i12 == 0.0 && i11 != i22) {
    const double c = cos(angle), s = sin(angle);
    const double g11p = qj * E * S[5], g22p = qj * E * S[4];
    gparam[3] = g11p * (c * c) + g22p * (s * s);
    gparam[4] = g11p * (s * s) + g22p * (c * c);
  } else {
    gparam[3] = nan;
    gparam[4] = nan;
  }
  gparam[5] = nan;
  gparam[6] = qj * G * S[3];
  gparam[7] = qj * G * A * S12;
  gparam[8] = nan;
  gparam[9] = nan;
  gparam[10] = nan;
}

}  // namespace fem
