// Time-stepping drop-in check: the reference's own step sequences, compiled
// unchanged from its headers and run on the host, next to the same sequences
// written with delfem2::hip types (include/delfem2_hip.hpp) on device arrays:
//   Newmark-beta  examples_oldgl_glfw/618_Fem3Tet/main.cpp:361-416
//                 (MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D, SetFixedBC,
//                 setRHS_Zero, CopyValue, Decompose, Solve_PCG, XPlusAYBZCW,
//                 XPlusAYBZ, XPlusAY)
//   stiffness warping  examples_oldgl_glfw/603_Fem3DStiffwarp/main.cpp:180-227
//                 (rotations, MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D, the
//                 same BC / ILU steps, Solve_PBiCGStab, XPlusAYBZ, XPlusAY)
// Reported: first-step assembly bit-identical, per-step iteration counts and
// the state after five steps.  Built by tests/cpp/Makefile where the reference
// tree exists; run by tests/test_gpu_dropin.py.
//
// Problem: 6x6x6 voxel cube x 6 tets, myu = 10, lambda = 1, rho = 1 (618's
// values), z = 0 clamped, g = (0, -3, -1), dt = 0.03.
#include <cmath>
#include <cstdio>
#include <vector>

#include "delfem2/ls_block_sparse.h"
#include "delfem2/ls_ilu_block_sparse.h"
#include "delfem2/fem_solidlinear.h"
#include "delfem2/lsitrsol.h"
#include "delfem2/mat3.h"
#include "delfem2/mat3_funcs.h"
#include "delfem2/svd3.h"
#include "delfem2/vec3.h"
#include "delfem2/vecxitrsol.h"
#include "delfem2/view_vectorx.h"
#include "delfem2_hip.hpp"

namespace dfm2 = delfem2;
namespace dh = delfem2::hip;

// 603_Fem3DStiffwarp/main.cpp:80-106 (an example-file function, the same
// steps over the reference's CVec3d / CMat3d / GetRotPolarDecomp)
static void RotationAtMeshPoints(std::vector<double> &aR, const std::vector<double> &aXYZ,
                                 const std::vector<double> &aDisp,
                                 const std::vector<unsigned int> &psup_ind,
                                 const std::vector<unsigned int> &psup) {
  const size_t np = aXYZ.size() / 3;
  aR.resize(np * 9);
  for (unsigned int ip = 0; ip < np; ++ip) {
    dfm2::CVec3d Pi(aXYZ[ip * 3 + 0], aXYZ[ip * 3 + 1], aXYZ[ip * 3 + 2]);
    dfm2::CVec3d pi(aXYZ[ip * 3 + 0] + aDisp[ip * 3 + 0], aXYZ[ip * 3 + 1] + aDisp[ip * 3 + 1],
                    aXYZ[ip * 3 + 2] + aDisp[ip * 3 + 2]);
    dfm2::CMat3d A;
    A.setZero();
    for (unsigned int jjp = psup_ind[ip]; jjp < psup_ind[ip + 1]; ++jjp) {
      const unsigned int jp = psup[jjp];
      dfm2::CVec3d Pj(aXYZ[jp * 3 + 0], aXYZ[jp * 3 + 1], aXYZ[jp * 3 + 2]);
      dfm2::CVec3d pj(aXYZ[jp * 3 + 0] + aDisp[jp * 3 + 0], aXYZ[jp * 3 + 1] + aDisp[jp * 3 + 1],
                      aXYZ[jp * 3 + 2] + aDisp[jp * 3 + 2]);
      A += dfm2::Mat3_OuterProduct(pj - pi, Pj - Pi);
    }
    dfm2::GetRotPolarDecomp(aR.data() + ip * 9, A.p_, 100);
  }
}

struct Dev {  // a device array of doubles
  explicit Dev(size_t n) : n(n) {
    dh::check(dfm2h_malloc(reinterpret_cast<void **>(&p), n * sizeof(double)));
    std::vector<double> z(n, 0.0);
    put(z);
  }
  ~Dev() { dfm2h_free(p); }
  void put(const std::vector<double> &h) { dh::check(dfm2h_memcpy_htod(p, h.data(), n * 8)); }
  std::vector<double> get() const {
    std::vector<double> h(n);
    dh::check(dfm2h_memcpy_dtoh(h.data(), p, n * 8));
    return h;
  }
  double *p = nullptr;
  size_t n;
};

static double rel(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / den);
}

int main() {
  const unsigned n = 6, m = n + 1, np = m * m * m, ndof = 3 * np;
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
  std::vector<int> flag(ndof, 0);
  for (unsigned ip = 0; ip < np; ++ip)
    if (xyz[3 * ip + 2] == 0.0) flag[3 * ip] = flag[3 * ip + 1] = flag[3 * ip + 2] = 1;
  const double myu = 10.0, lambda = 1.0, rho = 1.0, g[3] = {0.0, -3.0, -1.0}, dt = 0.03;
  const double gamma_nb = 1.0, beta_nb = 0.5, conv = 1.0e-10;
  const unsigned maxit = 1000, nstep = 5;

  dh::check(dfm2h_init(0));
  std::vector<unsigned> psup_ind(np + 1);
  size_t npsup = 0;
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), nullptr, 0, &npsup));
  std::vector<unsigned> psup(npsup);
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), psup.data(), npsup, &npsup));

  // ---- reference, host
  dfm2::CMatrixSparse<double> A;
  dfm2::CPreconditionerILU<double> ilu;
  A.Initialize(np, 3, true);
  A.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
  ilu.SetPattern0(A);
  // ---- device types
  dh::CMatrixSparse dA;
  dh::CPreconditionerILU dilu;
  dA.Initialize(np, 3, true);
  dA.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
  dilu.SetPattern0(dA);
  dh::TetMesh mesh(tet.data(), ntet, np);
  Dev dxyz(3 * np), dflag_store((ndof + 1) / 2), db(ndof), dx(ndof), dt0(ndof), dt1(ndof);
  dxyz.put(xyz);
  int *dflag = reinterpret_cast<int *>(dflag_store.p);
  dh::check(dfm2h_memcpy_htod(dflag, flag.data(), ndof * sizeof(int)));

  bool same_values = true, same_rhs = true;
  std::vector<int> it_ref[2], it_dev[2];
  double u_rel[2] = {0, 0}, v_rel[2] = {0, 0}, a_rel = 0;
  for (int scheme = 0; scheme < 2; ++scheme) {  // 0 Newmark-beta, 1 stiffness warping
    std::vector<double> aVal(ndof, 0.0), aVelo(ndof, 0.0), aAcc(ndof, 0.0), aR, vec_b;
    Dev du(ndof), dv(ndof), da(ndof), dR(9 * np);
    for (unsigned step = 0; step < nstep; ++step) {
      // -- reference
      A.setZero();
      vec_b.assign(ndof, 0.0);
      if (scheme == 0) {
        dfm2::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
            A, vec_b.data(), myu, lambda, rho, g, dt, gamma_nb, beta_nb, xyz.data(), np,
            tet.data(), ntet, aVal.data(), aVelo.data(), aAcc.data());
      } else {
        RotationAtMeshPoints(aR, xyz, aVal, psup_ind, psup);
        dfm2::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(A, vec_b.data(), myu, lambda, rho, g,
                                                          dt, xyz.data(), np, tet.data(),
                                                          ntet, aVal.data(), aVelo.data(), aR);
      }
      // -- device
      dA.setZero();
      dh::check(dfm2h_vec_set_zero(db.p, ndof));
      if (scheme == 0) {
        dh::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(dA, db.p, myu, lambda, rho, g, dt,
                                                          gamma_nb, beta_nb, dxyz.p, mesh, du.p,
                                                          dv.p, da.p);
      } else {
        dh::check(dfm2h_rotation_at_points(dA.h, dxyz.p, du.p, dR.p));
        dh::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(dA, db.p, myu, lambda, rho, g, dt,
                                                        dxyz.p, mesh, du.p, dv.p, dR.p);
      }
      if (step == 0) {  // same state (zero) and, for Newmark, the same arithmetic
        std::vector<double> vc(dA.ncrs() * 9), vd(9 * np);
        dA.download(vc.data(), vd.data());
        const std::vector<double> hb = db.get();
        if (scheme == 0) {
          same_values = same_values && vc == A.val_crs_ && vd == A.val_dia_;
          same_rhs = same_rhs && hb == vec_b;
        }
      }
      A.SetFixedBC(flag.data());
      dfm2::setRHS_Zero(vec_b, flag, 0);
      ilu.CopyValue(A);
      ilu.Decompose();
      std::vector<double> vec_x(ndof);
      dA.SetFixedBC(dflag);
      dh::setRHS_Zero(db.p, dflag, ndof, 0);
      dilu.CopyValue(dA);
      dilu.Decompose();
      if (scheme == 0) {
        std::vector<double> tmp0(ndof), tmp1(ndof);
        const auto h = dfm2::Solve_PCG(dfm2::ViewAsVectorXd(vec_b), dfm2::ViewAsVectorXd(vec_x),
                                       dfm2::ViewAsVectorXd(tmp0), dfm2::ViewAsVectorXd(tmp1),
                                       conv, maxit, A, ilu);
        it_ref[0].push_back(int(h.size()) - 1);
        const auto hd = Solve_PCG(dh::ViewAsVectorXd(db.p, ndof), dh::ViewAsVectorXd(dx.p, ndof),
                                  dh::ViewAsVectorXd(dt0.p, ndof),
                                  dh::ViewAsVectorXd(dt1.p, ndof), conv, maxit, dA, dilu);
        it_dev[0].push_back(int(hd.size()) - 1);
        dfm2::XPlusAYBZCW(aVal, ndof, flag, dt, aVelo, 0.5 * dt * dt, aAcc, dt * dt * beta_nb,
                          vec_x);
        dfm2::XPlusAYBZ(aVelo, ndof, flag, dt * gamma_nb, vec_x, dt, aAcc);
        dfm2::XPlusAY(aAcc, ndof, flag, 1.0, vec_x);
        dh::XPlusAYBZCW(du.p, ndof, dflag, dt, dv.p, 0.5 * dt * dt, da.p, dt * dt * beta_nb,
                        dx.p);
        dh::XPlusAYBZ(dv.p, ndof, dflag, dt * gamma_nb, dx.p, dt, da.p);
        dh::XPlusAY(da.p, ndof, dflag, 1.0, dx.p);
      } else {
        const auto h = dfm2::Solve_PBiCGStab(vec_b.data(), vec_x.data(), conv, maxit, A, ilu);
        it_ref[1].push_back(int(h.size()));
        const auto hd = dh::Solve_PBiCGStab(db.p, dx.p, conv, maxit, dA, dilu);
        it_dev[1].push_back(int(hd.size()));
        dfm2::XPlusAYBZ(aVal, ndof, flag, dt, vec_x, dt, aVelo);
        dfm2::XPlusAY(aVelo, ndof, flag, 1.0, vec_x);
        dh::XPlusAYBZ(du.p, ndof, dflag, dt, dx.p, dt, dv.p);
        dh::XPlusAY(dv.p, ndof, dflag, 1.0, dx.p);
      }
    }
    u_rel[scheme] = rel(du.get(), aVal);
    v_rel[scheme] = rel(dv.get(), aVelo);
    if (scheme == 0) a_rel = rel(da.get(), aAcc);
  }
  auto list = [](const std::vector<int> &v) {
    std::string s = "[";
    for (size_t k = 0; k < v.size(); ++k) s += (k ? ", " : "") + std::to_string(v[k]);
    return s + "]";
  };
  std::printf("{\"same_values\": %s, \"same_rhs\": %s, \"newmark_iterations_ref\": %s, "
              "\"newmark_iterations_dev\": %s, \"newmark_u_rel\": %.3e, \"newmark_v_rel\": %.3e, "
              "\"newmark_a_rel\": %.3e, \"stiffwarp_iterations_ref\": %s, "
              "\"stiffwarp_iterations_dev\": %s, \"stiffwarp_u_rel\": %.3e, "
              "\"stiffwarp_v_rel\": %.3e}\n",
              same_values ? "true" : "false", same_rhs ? "true" : "false",
              list(it_ref[0]).c_str(), list(it_dev[0]).c_str(), u_rel[0], v_rel[0], a_rel,
              list(it_ref[1]).c_str(), list(it_dev[1]).c_str(), u_rel[1], v_rel[1]);
  bool ok = same_values && same_rhs && u_rel[0] < 1e-7 && v_rel[0] < 1e-7 && a_rel < 1e-7 &&
            u_rel[1] < 1e-7 && v_rel[1] < 1e-7;
  for (int s = 0; s < 2; ++s)
    for (size_t k = 0; k < it_ref[s].size(); ++k)
      ok = ok && std::abs(it_ref[s][k] - it_dev[s][k]) <= 1;
  return ok ? 0 : 1;
}
