// rotation.hip -- per-node rotations of the stiffness-warping driver:
// 603_Fem3DStiffwarp/main.cpp:80-106 RotationAtMeshPoints over the matrix's
// one-ring pattern, with svd3.cpp:114-305 GetRotPolarDecomp (Jacobi sweeps
// on m^T m, eigenvalue sort, Svd3's U, R = U V^T) restated for the device,
// one thread per node.  The device's atan2 / cos / sin are not the host
// library's, so rotations agree with the reference to rounding, not bit for
// bit (sqrt and division are correctly rounded on both).
#include "handles.hpp"

namespace dfm2h {

__device__ inline void eigen_sym3(double u[9], const double sm[6], int nitr, double l[3]) {
  u[0] = u[4] = u[8] = 1.0;
  u[1] = u[2] = u[3] = u[5] = u[6] = u[7] = 0.0;
  l[0] = l[1] = l[2] = 0.0;
  const double dnrm = sm[0] * sm[0] + sm[1] * sm[1] + sm[2] * sm[2] +
                      2 * (sm[3] * sm[3] + sm[4] * sm[4] + sm[5] * sm[5]);
  if (dnrm < 1.0e-30) return;
  const double scale = sqrt(dnrm);
  const double invscl = 1.0 / scale;
  double sms[6] = {sm[0] * invscl, sm[1] * invscl, sm[2] * invscl,
                   sm[3] * invscl, sm[4] * invscl, sm[5] * invscl};
  for (int itr = 0; itr < nitr; itr++) {
    const double m0 = sms[0], m1 = sms[1], m2 = sms[2], m3 = sms[3], m4 = sms[4], m5 = sms[5];
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = u[k];
    const double a12 = fabs(m3), a20 = fabs(m4), a01 = fabs(m5);
    if (a12 >= a20 && a12 >= a01) {
      const double t = 0.5 * atan2(2 * m3, m2 - m1);
      const double ct = cos(t), st = sin(t);
      sms[1] = ct * ct * m1 + st * st * m2 - 2 * st * ct * m3;
      sms[2] = ct * ct * m2 + st * st * m1 + 2 * st * ct * m3;
      sms[3] = 0;
      sms[4] = st * m5 + ct * m4;
      sms[5] = ct * m5 - st * m4;
      u[1] = +ct * v[1] - st * v[2];
      u[2] = +st * v[1] + ct * v[2];
      u[4] = +ct * v[4] - st * v[5];
      u[5] = +st * v[4] + ct * v[5];
      u[7] = +ct * v[7] - st * v[8];
      u[8] = +st * v[7] + ct * v[8];
    } else if (a20 >= a01 && a20 >= a12) {
      const double t = 0.5 * atan2(2 * m4, m2 - m0);
      const double ct = cos(t), st = sin(t);
      sms[0] = ct * ct * m0 + st * st * m2 - 2 * st * ct * m4;
      sms[2] = ct * ct * m2 + st * st * m0 + 2 * st * ct * m4;
      sms[3] = st * m5 + ct * m3;
      sms[4] = 0;
      sms[5] = ct * m5 - st * m3;
      u[0] = +ct * v[0] - st * v[2];
      u[2] = +st * v[0] + ct * v[2];
      u[3] = +ct * v[3] - st * v[5];
      u[5] = +st * v[3] + ct * v[5];
      u[6] = +ct * v[6] - st * v[8];
      u[8] = +st * v[6] + ct * v[8];
    } else {
      const double t = 0.5 * atan2(2 * m5, m1 - m0);
      const double ct = cos(t), st = sin(t);
      sms[0] = ct * ct * m0 + st * st * m1 - 2 * st * ct * m5;
      sms[1] = ct * ct * m1 + st * st * m0 + 2 * st * ct * m5;
      sms[3] = st * m4 + ct * m3;
      sms[4] = ct * m4 - st * m3;
      sms[5] = 0;
      u[0] = +ct * v[0] - st * v[1];
      u[1] = +st * v[0] + ct * v[1];
      u[3] = +ct * v[3] - st * v[4];
      u[4] = +st * v[3] + ct * v[4];
      u[6] = +ct * v[6] - st * v[7];
      u[7] = +st * v[6] + ct * v[7];
    }
  }
  l[0] = scale * sms[0];
  l[1] = scale * sms[1];
  l[2] = scale * sms[2];
}

__device__ inline void swap_cols(double G[3], double V[9], int a, int b) {
  double t = G[a];
  G[a] = G[b];
  G[b] = t;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    t = V[3 * r + a];
    V[3 * r + a] = V[3 * r + b];
    V[3 * r + b] = t;
  }
}

__device__ inline void normalize3(double v[3]) {
  const double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= len;
  v[1] /= len;
  v[2] /= len;
}

__device__ inline void rot_polar_decomp(double R[9], const double m[9], int nitr) {
  const double mtm[6] = {m[0] * m[0] + m[3] * m[3] + m[6] * m[6],
                         m[1] * m[1] + m[4] * m[4] + m[7] * m[7],
                         m[2] * m[2] + m[5] * m[5] + m[8] * m[8],
                         m[1] * m[2] + m[4] * m[5] + m[7] * m[8],
                         m[2] * m[0] + m[5] * m[3] + m[8] * m[6],
                         m[0] * m[1] + m[3] * m[4] + m[6] * m[7]};
  double V[9], lv[3];
  eigen_sym3(V, mtm, nitr, lv);
  // svd3.cpp:60-108 SortEigen3
  if (lv[1] > lv[0]) swap_cols(lv, V, 0, 1);
  if (lv[2] > lv[1]) swap_cols(lv, V, 1, 2);
  if (lv[1] > lv[0]) swap_cols(lv, V, 0, 1);
  double u0[3], u1[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    u0[r] = m[3 * r + 0] * V[0] + m[3 * r + 1] * V[3] + m[3 * r + 2] * V[6];
    u1[r] = m[3 * r + 0] * V[1] + m[3 * r + 1] * V[4] + m[3 * r + 2] * V[7];
  }
  const double det = +V[0] * V[4] * V[8] + V[3] * V[7] * V[2] + V[6] * V[1] * V[5] -
                     V[0] * V[7] * V[5] - V[6] * V[4] * V[2] - V[3] * V[1] * V[8];
  if (det < 0) {
    V[2] *= -1;
    V[5] *= -1;
    V[8] *= -1;
  }
  double u2[3];
  const double sql0 = u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2];
  if (sql0 > 1.0e-20) {
    normalize3(u0);
    const double sql1 = u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2];
    if (sql1 < 1.0e-20) {
      u1[0] = 1.0 - fabs(u0[0]);
      u1[1] = 1.0 - fabs(u0[1]);
      u1[2] = 1.0 - fabs(u0[2]);
    } else {
      normalize3(u1);
    }
    const double d01 = u0[0] * u1[0] + u0[1] * u1[1] + u0[2] * u1[2];
    u1[0] -= d01 * u0[0];
    u1[1] -= d01 * u0[1];
    u1[2] -= d01 * u0[2];
    normalize3(u1);
    u2[0] = u0[1] * u1[2] - u1[1] * u0[2];
    u2[1] = u0[2] * u1[0] - u1[2] * u0[0];
    u2[2] = u0[0] * u1[1] - u1[0] * u0[1];
  } else {
    u0[0] = 1; u0[1] = 0; u0[2] = 0;
    u1[0] = 0; u1[1] = 1; u1[2] = 0;
    u2[0] = 0; u2[1] = 0; u2[2] = 1;
  }
  const double U[9] = {u0[0], u1[0], u2[0], u0[1], u1[1], u2[1], u0[2], u1[2], u2[2]};
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      R[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] +
                     U[i * 3 + 2] * V[j * 3 + 2];
}

__global__ void __launch_bounds__(128)
k_rotation_at_points(unsigned np, const unsigned *__restrict__ colind,
                     const unsigned *__restrict__ rowptr, const double *__restrict__ xyz,
                     const double *__restrict__ disp, double *__restrict__ R) {
  const unsigned ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double Pi[3], pi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Pi[k] = xyz[(size_t)ip * 3 + k];
    pi[k] = xyz[(size_t)ip * 3 + k] + (disp ? disp[(size_t)ip * 3 + k] : 0.0);
  }
  double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (unsigned jj = colind[ip]; jj < colind[ip + 1]; ++jj) {
    const unsigned jp = rowptr[jj];
    double a[3], b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double Pj = xyz[(size_t)jp * 3 + k];
      const double pj = xyz[(size_t)jp * 3 + k] + (disp ? disp[(size_t)jp * 3 + k] : 0.0);
      a[k] = pj - pi[k];
      b[k] = Pj - Pi[k];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) A[i * 3 + j] += a[i] * b[j];
  }
  double Rm[9];
  rot_polar_decomp(Rm, A, 100);
#pragma unroll
  for (int k = 0; k < 9; ++k) R[(size_t)ip * 9 + k] = Rm[k];
}

}  // namespace dfm2h

using namespace dfm2h;

extern "C" {

int dfm2h_rotation_at_points(dfm2h_bsr A, const double *xyz, const double *disp, double *R) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A && xyz && R, DFM2H_ERR_INVALID, "null argument");
    DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_INVALID, "matrix has no pattern");
    DFM2H_DEVICE_ARG(xyz);
    DFM2H_DEVICE_ARG(disp);
    DFM2H_DEVICE_ARG(R);
    if (!A->nblk) return;
    hipLaunchKernelGGL(k_rotation_at_points, dim3(ceil_div(A->nblk, 128)), dim3(128), 0,
                       stream(), (unsigned)A->nblk, A->pat.colind_d.p, A->pat.rowptr_d.p, xyz,
                       disp, R);
    check_launch("k_rotation_at_points");
  });
}

}  // extern "C"
