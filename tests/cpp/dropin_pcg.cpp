// Drop-in check: the REFERENCE's own Solve_PCG template (lsitrsol.h:80-148),
// compiled unchanged, instantiated on the device types of delfem2_hip.hpp,
// next to the fused device overload found by ADL.  Built by tests/cpp/Makefile
// where the reference tree exists; run by tests/test_gpu_dropin.py.
//
// Problem: 10^3 voxel cube x 6 tets (3,993 DOF), E=1 (myu=1/2.6,
// lambda=0.3/(1.3*0.4)), z=0 clamped, body force (0,0,-1e-3).
#include <cmath>
#include <cstdio>
#include <vector>

#include "delfem2/lsitrsol.h"  // the reference template, unchanged
#include "delfem2_hip.hpp"

namespace dh = delfem2::hip;

template <class T>
static T *to_dev(const std::vector<T> &v) {
  void *p = nullptr;
  dh::check(dfm2h_malloc(&p, sizeof(T) * v.size()));
  dh::check(dfm2h_memcpy_htod(p, v.data(), sizeof(T) * v.size()));
  return static_cast<T *>(p);
}

int main() {
  const unsigned n = 10, m = n + 1, np = m * m * m;
  std::vector<double> xyz(3 * np);
  for (unsigned ix = 0; ix < m; ++ix)
    for (unsigned iy = 0; iy < m; ++iy)
      for (unsigned iz = 0; iz < m; ++iz) {
        const unsigned ip = (ix * m + iy) * m + iz;
        xyz[3 * ip + 0] = double(ix) / n;
        xyz[3 * ip + 1] = double(iy) / n;
        xyz[3 * ip + 2] = double(iz) / n;
      }
  const unsigned off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const unsigned split[6][4] = {{0, 1, 2, 6}, {0, 1, 6, 5}, {0, 4, 5, 6},
                                {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 4, 6, 7}};
  std::vector<unsigned> tet;
  for (unsigned ix = 0; ix < n; ++ix)
    for (unsigned iy = 0; iy < n; ++iy)
      for (unsigned iz = 0; iz < n; ++iz) {
        unsigned c[8];
        for (int k = 0; k < 8; ++k)
          c[k] = ((ix + off[k][0]) * m + iy + off[k][1]) * m + iz + off[k][2];
        for (auto &s : split)
          for (int k = 0; k < 4; ++k) tet.push_back(c[s[k]]);
      }
  const size_t ntet = tet.size() / 4, ndof = 3 * np;
  std::vector<int> flag(ndof, 0);
  for (unsigned ip = 0; ip < np; ++ip)
    if (xyz[3 * ip + 2] == 0.0) flag[3 * ip] = flag[3 * ip + 1] = flag[3 * ip + 2] = 1;

  dh::check(dfm2h_init(0));
  std::vector<unsigned> psup_ind(np + 1);
  size_t npsup = 0;
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), nullptr, 0, &npsup));
  std::vector<unsigned> psup(npsup);
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), psup.data(), npsup, &npsup));

  dh::CMatrixSparse A;
  A.Initialize(np, 3, true);
  A.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
  dh::CPreconditionerILU ilu;
  ilu.SetPattern0(A);
  dh::TetMesh mesh(tet.data(), ntet, np);
  double *d_xyz = to_dev(xyz);
  int *d_flag = to_dev(flag);
  double *d_b = to_dev(std::vector<double>(ndof, 0.0));
  double *d_r = to_dev(std::vector<double>(ndof, 0.0));
  double *d_x = to_dev(std::vector<double>(ndof, 0.0));
  double *d_t0 = to_dev(std::vector<double>(ndof, 0.0));
  double *d_t1 = to_dev(std::vector<double>(ndof, 0.0));
  const double E = 1.0, nu = 0.3, g[3] = {0.0, 0.0, -1e-3};
  const double myu = E / (2.0 * (1.0 + nu)), lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  A.setZero();
  if (!dh::MergeLinSys_SolidLinear_Static_MeshTet3D(A, d_b, myu, lambda, 1.0, g, d_xyz, mesh,
                                                    nullptr))
    return 3;
  A.SetFixedBC(d_flag);
  dh::setRHS_Zero(d_b, d_flag, ndof, 0);
  ilu.CopyValue(A);
  if (!ilu.Decompose()) return 4;

  // (1) the reference template, unchanged, on device views
  dh::check(dfm2h_memcpy_dtod(d_r, d_b, sizeof(double) * ndof));
  std::vector<double> h1 = delfem2::Solve_PCG(
      dh::ViewAsVectorXd(d_r, ndof), dh::ViewAsVectorXd(d_x, ndof),
      dh::ViewAsVectorXd(d_t0, ndof), dh::ViewAsVectorXd(d_t1, ndof), 1e-8, 1000, A, ilu);
  std::vector<double> x1(ndof);
  dh::check(dfm2h_memcpy_dtoh(x1.data(), d_x, sizeof(double) * ndof));
  // (2) unqualified call: ADL picks the fused device overload
  dh::check(dfm2h_memcpy_dtod(d_r, d_b, sizeof(double) * ndof));
  std::vector<double> h2 = Solve_PCG(dh::ViewAsVectorXd(d_r, ndof), dh::ViewAsVectorXd(d_x, ndof),
                                     dh::ViewAsVectorXd(d_t0, ndof),
                                     dh::ViewAsVectorXd(d_t1, ndof), 1e-8, 1000, A, ilu);
  std::vector<double> x2(ndof);
  dh::check(dfm2h_memcpy_dtoh(x2.data(), d_x, sizeof(double) * ndof));
  double num = 0.0, den = 0.0, umax = 0.0;
  for (size_t i = 0; i < ndof; ++i) {
    num += (x1[i] - x2[i]) * (x1[i] - x2[i]);
    den += x1[i] * x1[i];
    umax = std::fmax(umax, std::fabs(x1[i]));
  }
  const double rel = std::sqrt(num / den);
  std::printf("{\"template_iterations\": %zu, \"fused_iterations\": %zu, \"x_rel_diff\": %.3e, "
              "\"max_abs_u\": %.9e}\n", h1.size() - 1, h2.size() - 1, rel, umax);
  for (void *p : {(void *)d_xyz, (void *)d_flag, (void *)d_b, (void *)d_r, (void *)d_x,
                  (void *)d_t0, (void *)d_t1})
    dfm2h_free(p);
  const bool ok = rel < 1e-9 && (h1.size() == h2.size() || h1.size() + 1 == h2.size() ||
                                 h2.size() + 1 == h1.size());
  return ok ? 0 : 1;
}
