// forces.hip -- element post-processing on the device
// (extract_element_analysis_result: truss.rs:281-333, beam.rs:803-993,
// plate.rs:1196-1409, methods_for_element_analysis.rs:27-58), one thread per
// element, from global displacements (6 per node).
//   truss: ForceR
//   beam:  ForceR, ForceS, ForceT, MomentR, MomentS (node 1, average, node 2),
//          MomentT (node 1, average, node 2)
//   plate: MembraneForceR, S, RS, BendingMomentR, S, RS, ShearForceRT, ST,
//          from B summed over the four nodal points
// The element frames are the ones the assembly set up (R, beam principal
// moments), so the displacements go to the same local frames.
#include <hip/hip_runtime.h>

#include <string>

#include "device_state.hpp"
#include "element_math.hpp"
#include "fem_device.hpp"

namespace femdev {

namespace {

// local = R g over `blocks` blocks of three
__device__ inline void to_local(const double* __restrict__ R, const double* g, int blocks, double* loc) {
  for (int b = 0; b < blocks; ++b)
    for (int i = 0; i < 3; ++i)
      loc[3 * b + i] = (R[3 * i] * g[3 * b] + R[3 * i + 1] * g[3 * b + 1]) + R[3 * i + 2] * g[3 * b + 2];
}

__global__ void truss_forces_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ prm,
                                    const double* __restrict__ Rs, const double* __restrict__ xyz,
                                    const double* __restrict__ u, size_t n, double* __restrict__ out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t a = conn[2 * e], b = conn[2 * e + 1];
  double g[6], loc[6];
  for (int i = 0; i < 3; ++i) {
    g[i] = u[6 * size_t(a) + i];
    g[3 + i] = u[6 * size_t(b) + i];
  }
  to_local(Rs + 9 * e, g, 2, loc);
  const double dx = xyz[3 * size_t(b)] - xyz[3 * size_t(a)], dy = xyz[3 * size_t(b) + 1] - xyz[3 * size_t(a) + 1],
               dz = xyz[3 * size_t(b) + 2] - xyz[3 * size_t(a) + 2];
  const double inv_j = 1.0 / fem::bar_jacobian(fem::norm3(dx, dy, dz));
  const double E = prm[3 * e], A = prm[3 * e + 1], A2 = prm[3 * e + 2];
  const double area = isnan(A2) ? A : (A2 - A) / 2.0 * 0.0 + A - (A2 - A) / 2.0 * -1.0;  // truss.rs:120-130, r = 0
  const double strain = (-0.5 * inv_j) * loc[0] + (0.5 * inv_j) * loc[3];
  out[e] = strain * (E * area / 1.0);
}

__global__ void beam_forces_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ prm,
                                   const double* __restrict__ Rs, const double* __restrict__ derived,
                                   const double* __restrict__ xyz, const double* __restrict__ u, size_t n,
                                   double* __restrict__ out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t a = conn[2 * e], b = conn[2 * e + 1];
  double g[12], loc[12];
  for (int i = 0; i < 6; ++i) {
    g[i] = u[6 * size_t(a) + i];
    g[6 + i] = u[6 * size_t(b) + i];
  }
  to_local(Rs + 9 * e, g, 4, loc);
  const double dx = xyz[3 * size_t(b)] - xyz[3 * size_t(a)], dy = xyz[3 * size_t(b) + 1] - xyz[3 * size_t(a) + 1],
               dz = xyz[3 * size_t(b) + 2] - xyz[3 * size_t(a) + 2];
  const double L = fem::norm3(dx, dy, dz);
  const double inv_j = 1.0 / fem::bar_jacobian(L);
  const double* p = prm + 11 * e;
  const double E = p[0], nu = p[1], A = p[2], it = p[6], k = p[7];
  const double i11p = derived[2 * e], i22p = derived[2 * e + 1];
  const double G = E / (2.0 * (1.0 + nu));
  const double d1 = -0.5 * inv_j, d2 = 0.5 * inv_j;  // dh/dx at r = 0; h1 = h2 = 1/2
  const double fr = (d1 * loc[0] + d2 * loc[6]) * (E * A / 1.0);
  const double fs = (((d1 * loc[1] + d2 * loc[7]) - 0.5 * loc[5]) - 0.5 * loc[11]) * (G * A * k / 1.0);
  const double ft = (((d1 * loc[2] + d2 * loc[8]) - 0.5 * loc[4]) - 0.5 * loc[10]) * (G * A * k / 1.0);
  const double mr = (d1 * loc[3] + d2 * loc[9]) * (G * it / 1.0);
  const double ms = (d1 * loc[4] + d2 * loc[10]) * (E * i22p / 1.0);
  const double mt = (d1 * loc[5] + d2 * loc[11]) * (E * i11p / 1.0);
  double* o = out + 10 * e;
  o[0] = fr;
  o[1] = fs;
  o[2] = ft;
  o[3] = mr;
  o[4] = ms + L * ft / 2.0;
  o[5] = ms;
  o[6] = ms - L * ft / 2.0;
  o[7] = mt + L * fs / 2.0;
  o[8] = mt;
  o[9] = mt - L * fs / 2.0;
}

__global__ __launch_bounds__(256) void plate_forces_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ prm,
                                    const double* __restrict__ Rs, const double* __restrict__ xyz,
                                    const double* __restrict__ u, size_t n, double* __restrict__ out) {
  const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* R = Rs + 9 * e;
  uint32_t nd[4];
  double g[24], loc[24], X[4][3];
  for (int a = 0; a < 4; ++a) {
    nd[a] = conn[4 * e + a];
    for (int i = 0; i < 3; ++i) X[a][i] = xyz[3 * size_t(nd[a]) + i];
    for (int i = 0; i < 6; ++i) g[6 * a + i] = u[6 * size_t(nd[a]) + i];
  }
  to_local(R, g, 8, loc);
  // local coordinates of nodes 1, 2, 4 relative to node 3 (quad:340-378)
  double xl[6];
  const int others[3] = {0, 1, 3};
  for (int m = 0; m < 3; ++m) {
    const int a = others[m];
    const double r0 = X[a][0] - X[2][0], r1 = X[a][1] - X[2][1], r2 = X[a][2] - X[2][2];
    xl[2 * m] = (R[0] * r0 + R[1] * r1) + R[2] * r2;
    xl[2 * m + 1] = (R[3] * r0 + R[4] * r1) + R[5] * r2;
  }
  double em[3] = {0.0, 0.0, 0.0}, eb[3] = {0.0, 0.0, 0.0}, es[2] = {0.0, 0.0};
  const double pts[4][2] = {{1.0, 1.0}, {-1.0, 1.0}, {-1.0, -1.0}, {1.0, -1.0}};
  for (int k = 0; k < 4; ++k) {
    fem::PlateIP ip;
    fem::plate_ip(xl, pts[k][0], pts[k][1], ip);
    for (int a = 0; a < 4; ++a) {
      const double* l = loc + 6 * a;
      em[0] += ip.dx[a] * l[0];
      em[1] += ip.dy[a] * l[1];
      em[2] += ip.dy[a] * l[0] + ip.dx[a] * l[1];
      eb[0] += -ip.dx[a] * l[4];
      eb[1] += ip.dy[a] * l[3];
      eb[2] += ip.dx[a] * l[3] - ip.dy[a] * l[4];
      es[0] += (ip.sr[a][0] * l[2] + ip.sr[a][1] * l[3]) + ip.sr[a][2] * l[4];
      es[1] += (ip.ss[a][0] * l[2] + ip.ss[a][1] * l[3]) + ip.ss[a][2] * l[4];
    }
  }
  const double E = prm[4 * e], nu = prm[4 * e + 1], t = prm[4 * e + 2], ks = prm[4 * e + 3];
  const double base[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1.0 - nu) / 2.0};
  const double cmm = E / (1.0 - nu * nu), cbm = E * t / (2.0 * (1.0 - nu * nu)), csm = E / (2.0 * (1.0 + nu));
  double fm[3], fb[3];
  for (int i = 0; i < 3; ++i) {
    double am = 0.0, ab = 0.0;
    for (int j = 0; j < 3; ++j) {
      am += (base[3 * i + j] * cmm) * em[j];
      ab += (base[3 * i + j] * cbm) * eb[j];
    }
    fm[i] = am * (t / 4.0);
    fb[i] = ab * (t * t / 24.0);
  }
  double* o = out + 8 * e;
  o[0] = fm[0];
  o[1] = fm[1];
  o[2] = fm[2];
  o[3] = fb[1];
  o[4] = fb[0];
  o[5] = fb[2];
  o[6] = (csm * es[0]) * (t * ks / 4.0);
  o[7] = (csm * es[1]) * (t * ks / 4.0);
}

// u = imposed displacements, then u[k_aa_indexes] = u_a
__global__ void copy_kernel(const double* __restrict__ src, int64_t n, double* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void scatter_ua_kernel(const int64_t* __restrict__ aa, const double* __restrict__ ua, int64_t na,
                                  double* __restrict__ u) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < na) u[aa[i]] = ua[i];
}

// beam.rs:775-797: one point (r = 0, alpha = 2), h1 = h2 = 1/2
__global__ void line_load_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ xyz,
                                 const uint32_t* __restrict__ pos, const double* __restrict__ q, size_t n,
                                 double* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t e = pos[i];
  const uint32_t a = conn[2 * e], b = conn[2 * e + 1];
  const double L = fem::norm3(xyz[3 * size_t(b)] - xyz[3 * size_t(a)], xyz[3 * size_t(b) + 1] - xyz[3 * size_t(a) + 1],
                              xyz[3 * size_t(b) + 2] - xyz[3 * size_t(a) + 2]);
  const double f = 0.0 + (0.5 * q[i]) * (fem::bar_jacobian(L) * 2.0);
  out[2 * i] = f;
  out[2 * i + 1] = f;
}

// plate.rs:1145-1185: h_a(r, s) p det J(r, s), summed over the four points
__global__ void surface_load_kernel(const uint32_t* __restrict__ conn, const double* __restrict__ Rs,
                                    const double* __restrict__ xyz, const uint32_t* __restrict__ pos,
                                    const double* __restrict__ q, size_t n, double* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t e = pos[i];
  const double* R = Rs + 9 * e;
  double X[4][3];
  for (int a = 0; a < 4; ++a)
    for (int k = 0; k < 3; ++k) X[a][k] = xyz[3 * size_t(conn[4 * e + a]) + k];
  double xl[6];
  const int others[3] = {0, 1, 3};
  for (int m = 0; m < 3; ++m) {
    const int a = others[m];
    const double r0 = X[a][0] - X[2][0], r1 = X[a][1] - X[2][1], r2 = X[a][2] - X[2][2];
    xl[2 * m] = (R[0] * r0 + R[1] * r1) + R[2] * r2;
    xl[2 * m + 1] = (R[3] * r0 + R[4] * r1) + R[5] * r2;
  }
  const double g = fem::plate_gauss();
  double f[4] = {0.0, 0.0, 0.0, 0.0};
  for (int ip = 0; ip < 4; ++ip) {
    fem::PlatePoint pp;
    fem::plate_point(xl, ip, pp);
    const double r = (ip == 0 || ip == 3) ? g : -g, s = ip < 2 ? g : -g;
    const double h[4] = {0.25 * (1.0 + r) * (1.0 + s), 0.25 * (1.0 - r) * (1.0 + s), 0.25 * (1.0 - r) * (1.0 - s),
                         0.25 * (1.0 + r) * (1.0 - s)};
    for (int a = 0; a < 4; ++a) f[a] = f[a] + (h[a] * q[i]) * (pp.w * 1.0);
  }
  for (int a = 0; a < 4; ++a) out[4 * i + a] = f[a];
}

}  // namespace

std::string nodal_loads(Device* d, int type, const uint32_t* pos, const double* values, size_t n, double* out) {
  if (type != 1 && type != 2) return "distributed loads apply to beams (line) and plates (surface)";
  if (!n) return "";
  const int per = type == 1 ? 2 : 4;
  uint32_t* dpos = nullptr;
  double *dval = nullptr, *dout = nullptr;
  std::string e;
  auto cleanup = [&]() {
    if (dpos) (void)hipFree(dpos);
    if (dval) (void)hipFree(dval);
    if (dout) (void)hipFree(dout);
  };
  if ((e = dupload(dpos, pos, n)).empty() && (e = dupload(dval, values, n)).empty() &&
      (e = dalloc(dout, n * size_t(per))).empty()) {
    const unsigned B = 256, G = unsigned((n + B - 1) / B);
    if (type == 1)
      line_load_kernel<<<G, B>>>(d->conn[1], d->xyz, dpos, dval, n, dout);
    else
      surface_load_kernel<<<G, B>>>(d->conn[2], d->R[2], d->xyz, dpos, dval, n, dout);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(out, dout, sizeof(double) * n * size_t(per), hipMemcpyDeviceToHost);
    if (err != hipSuccess) e = std::string("nodal load kernels: ") + hipGetErrorString(err);
  }
  cleanup();
  return e;
}

std::string solution_displacements(Device* d, void* stream, const double** u) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!d->sol_done) return "no solution computed";
  const int64_t n = d->n_rows;
  if (d->sep_bc_rows != n) return "boundary conditions are not uploaded";
  if (d->ucap < size_t(n) || !d->u_glob) {
    auto e = dalloc(d->u_glob, size_t(n ? n : 1));
    if (!e.empty()) return e;
    d->ucap = size_t(n ? n : 1);
  }
  const unsigned B = 256;
  if (n) {
    copy_kernel<<<unsigned((n + B - 1) / B), B, 0, s>>>(d->sep_disp, n, d->u_glob);
    HIP_TRY(hipGetLastError());
  }
  if (d->sol_na) {
    scatter_ua_kernel<<<unsigned((d->sol_na + B - 1) / B), B, 0, s>>>(d->sep_aa, d->sol_vec[1], d->sol_na, d->u_glob);
    HIP_TRY(hipGetLastError());
  }
  *u = d->u_glob;
  return "";
}

std::string element_forces(Device* d, const double* u, double* const out[3], void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned B = 256;
  const size_t nt = d->n_elem[0], nbm = d->n_elem[1], np = d->n_elem[2];
  if (nt && out[0]) {
    truss_forces_kernel<<<unsigned((nt + B - 1) / B), B, 0, s>>>(d->conn[0], d->params[0], d->R[0], d->xyz, u, nt,
                                                                 out[0]);
    HIP_TRY(hipGetLastError());
  }
  if (nbm && out[1]) {
    beam_forces_kernel<<<unsigned((nbm + B - 1) / B), B, 0, s>>>(d->conn[1], d->params[1], d->R[1], d->derived,
                                                                 d->xyz, u, nbm, out[1]);
    HIP_TRY(hipGetLastError());
  }
  if (np && out[2]) {
    plate_forces_kernel<<<unsigned((np + B - 1) / B), B, 0, s>>>(d->conn[2], d->params[2], d->R[2], d->xyz, u, np,
                                                                 out[2]);
    HIP_TRY(hipGetLastError());
  }
  return "";
}

}  // namespace femdev
