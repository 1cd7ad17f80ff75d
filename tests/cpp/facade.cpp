// Facade check: the REFERENCE's LinearSystemSolver_BlockSparseILU
// (ls_solver_block_sparse_ilu.h:18-103, compiled unchanged, host) and
// delfem2::hip::LinearSystemSolver_BlockSparseILU (include/delfem2_hip.hpp,
// device solves) driven by the same caller code: element matrices from the
// reference's EMat_SolidLinear_Static_Tet merged element by element, a value
// added to a diagonal, fixed BCs, then Solve_PcgIlu and Solve_Cg on both.
// Built by tests/cpp/Makefile where the reference tree exists; run by
// tests/test_gpu_dropin.py.
//
// Problem: 8x8x8 voxel cube x 6 tets, E=1, nu=0.3, z=0 clamped, body force
// (0,0,-1e-3).
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

// ls_block_sparse.h first: fem_solidlinear.h's drivers call Merge<> and only
// include it themselves in static-library mode
#include "delfem2/ls_block_sparse.h"
#include "delfem2/fem_solidlinear.h"
#include "delfem2/ls_solver_block_sparse_ilu.h"
#include "delfem2/vecxitrsol.h"
#include "delfem2_hip.hpp"

namespace dfm2 = delfem2;
namespace dh = delfem2::hip;

template <class F>
static void assemble(F &&merge, std::vector<double> &vec_r, const std::vector<double> &xyz,
                     const std::vector<unsigned> &tet) {
  const double E = 1.0, nu = 0.3, g[3] = {0.0, 0.0, -1e-3};
  const double myu = E / (2.0 * (1.0 + nu)), lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  for (size_t e = 0; e < tet.size() / 4; ++e) {
    const unsigned *ip = &tet[4 * e];
    double P[4][3], U[4][3] = {}, emat[4][4][3][3] = {}, eres[4][3];
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) P[k][d] = xyz[3 * ip[k] + d];
    const double vol = dfm2::femutil::TetVolume3D(P[0], P[1], P[2], P[3]);
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) eres[k][d] = vol * g[d] * 0.25;
    dfm2::EMat_SolidLinear_Static_Tet(emat, eres, myu, lambda, P, U, true);
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) vec_r[3 * ip[k] + d] += eres[k][d];
    merge(ip, emat);
  }
}

static double rel(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / den);
}

int main() {
  const unsigned n = 8, m = n + 1, np = m * m * m;
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
  const size_t ntet = tet.size() / 4;
  dh::check(dfm2h_init(0));
  std::vector<unsigned> psup_ind(np + 1);
  size_t npsup = 0;
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), nullptr, 0, &npsup));
  std::vector<unsigned> psup(npsup);
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), psup.data(), npsup, &npsup));

  dfm2::LinearSystemSolver_BlockSparseILU ref;
  dh::LinearSystemSolver_BlockSparseILU dev;
  ref.Initialize(np, 3, psup_ind, psup);
  dev.Initialize(np, 3, psup_ind, psup);
  auto setup = [&](auto &ls) {
    ls.BeginMerge();
    assemble([&](const unsigned *ip, const double (*emat)[4][3][3]) {
      ls.template Merge<4, 4, 3, 3>(ip, ip, emat);
    }, ls.vec_r, xyz, tet);
    ls.AddValueToDiagonal(7, 1, 0.25);
    for (unsigned ip = 0; ip < np; ++ip)
      if (xyz[3 * ip + 2] == 0.0) ls.dof_bcflag[3 * ip] = ls.dof_bcflag[3 * ip + 1] =
          ls.dof_bcflag[3 * ip + 2] = 1;
  };
  setup(ref);
  setup(dev);
  // merged values identical bit for bit before the solve
  bool same_vals = ref.matrix.val_crs_ == dev.val_crs_ && ref.matrix.val_dia_ == dev.val_dia_;
  bool same_rhs = ref.vec_r == dev.vec_r;
  ref.Solve_PcgIlu();
  dev.Solve_PcgIlu();
  const double rel_pcg = rel(dev.vec_x, ref.vec_x);
  const size_t it_pcg = dev.conv_hist.size() - 1;
  // Solve_Cg after re-merging (the solve applied the BCs to the matrix)
  setup(ref);
  setup(dev);
  ref.Solve_Cg();
  dev.Solve_Cg();
  const double rel_cg = rel(dev.vec_x, ref.vec_x);
  std::printf("{\"same_values\": %s, \"same_rhs\": %s, \"pcg_iterations\": %zu, \"pcg_x_rel\": %.3e, "
              "\"cg_iterations_ref\": %zu, \"cg_iterations_dev\": %zu, \"cg_x_rel\": %.3e}\n",
              same_vals ? "true" : "false", same_rhs ? "true" : "false", it_pcg, rel_pcg,
              ref.conv_hist.size(), dev.conv_hist.size(), rel_cg);
  const bool ok = same_vals && same_rhs && rel_pcg < 1e-9 && rel_cg < 1e-8 &&
                  (ref.conv_hist.size() + 1 >= dev.conv_hist.size()) &&
                  (dev.conv_hist.size() + 1 >= ref.conv_hist.size());
  return ok ? 0 : 1;
}
