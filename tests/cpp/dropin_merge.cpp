// Drop-in check of the per-element Merge: the REFERENCE's own element-loop
// templates (fem_solidlinear.h:192-237 static, :331-379 Newmark-beta,
// :381-443 backward Euler, :445-530 stiffness-warped backward Euler),
// compiled unchanged, instantiated once on the reference's host
// CMatrixSparse<double> and once on delfem2::hip::CMatrixSparse, where their
// unqualified Merge<4,4,3,3,double> finds the device Merge by ADL.  The device
// values must equal the host ones bit for bit, and so must the device bulk
// assembly of the same kind.  Also checks that a host array passed where a
// device pointer is required is refused with DFM2H_ERR_INVALID.  Built by
// tests/cpp/Makefile where the reference tree exists; run by
// tests/test_gpu_dropin.py.
//
// Problem: 9 x 7 x 6 voxel grid x 6 tets, seeded displacement / velocity /
// acceleration / rotation fields, myu = 0.4, lambda = 0.9, rho = 1.3.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "delfem2_hip.hpp"  // first: it makes Merge a visible template name
#include "delfem2/fem_solidlinear.h"  // the reference templates, unchanged
#include "delfem2/ls_block_sparse.h"

namespace dh = delfem2::hip;

template <class T>
static T *to_dev(const std::vector<T> &v) {
  void *p = nullptr;
  dh::check(dfm2h_malloc(&p, sizeof(T) * v.size()));
  dh::check(dfm2h_memcpy_htod(p, v.data(), sizeof(T) * v.size()));
  return static_cast<T *>(p);
}

template <class T>
static std::vector<T> to_host(const T *p, size_t n) {
  std::vector<T> v(n);
  dh::check(dfm2h_memcpy_dtoh(v.data(), p, sizeof(T) * n));
  return v;
}

static bool same_bits(const std::vector<double> &a, const std::vector<double> &b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(double) * a.size()) == 0;
}

// seeded values in [-s, s) (SplitMix64)
static std::vector<double> field(size_t n, uint64_t seed, double s) {
  std::vector<double> v(n);
  for (auto &x : v) {
    uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    x = s * (2.0 * double(z >> 11) / 9007199254740992.0 - 1.0);
  }
  return v;
}

int main() {
  const unsigned nx = 9, ny = 7, nz = 6, mx = nx + 1, my = ny + 1, mz = nz + 1;
  const unsigned np = mx * my * mz;
  std::vector<double> xyz(3 * np);
  for (unsigned ix = 0; ix < mx; ++ix)
    for (unsigned iy = 0; iy < my; ++iy)
      for (unsigned iz = 0; iz < mz; ++iz) {
        const unsigned ip = (ix * my + iy) * mz + iz;
        xyz[3 * ip + 0] = double(ix) / nx;
        xyz[3 * ip + 1] = double(iy) / nx;
        xyz[3 * ip + 2] = double(iz) / nx;
      }
  const unsigned off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const unsigned split[6][4] = {{0, 1, 2, 6}, {0, 1, 6, 5}, {0, 4, 5, 6},
                                {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 4, 6, 7}};
  std::vector<unsigned> tet;
  for (unsigned ix = 0; ix < nx; ++ix)
    for (unsigned iy = 0; iy < ny; ++iy)
      for (unsigned iz = 0; iz < nz; ++iz) {
        unsigned c[8];
        for (int k = 0; k < 8; ++k)
          c[k] = ((ix + off[k][0]) * my + iy + off[k][1]) * mz + iz + off[k][2];
        for (auto &s : split)
          for (int k = 0; k < 4; ++k) tet.push_back(c[s[k]]);
      }
  const size_t ntet = tet.size() / 4, ndof = 3 * np;
  const std::vector<double> u = field(ndof, 11, 1e-2), v = field(ndof, 12, 1e-1),
                            a = field(ndof, 13, 1.0);
  // near-identity rotations (their exact value does not matter here)
  std::vector<double> R(9 * np);
  {
    const std::vector<double> w = field(3 * np, 14, 0.2);
    for (unsigned ip = 0; ip < np; ++ip) {
      const double c = std::cos(w[3 * ip]), s = std::sin(w[3 * ip]);
      const double r[9] = {c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0};
      std::copy(r, r + 9, R.begin() + 9 * ip);
    }
  }
  const double myu = 0.4, lambda = 0.9, rho = 1.3, g[3] = {0.1, -0.2, -1.0};
  const double dt = 0.03, gamma_nb = 0.6, beta_nb = 0.36;

  dh::check(dfm2h_init(0));
  std::vector<unsigned> psup_ind(np + 1);
  size_t npsup = 0;
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), nullptr, 0, &npsup));
  std::vector<unsigned> psup(npsup);
  dh::check(dfm2h_psup_tet(tet.data(), ntet, np, psup_ind.data(), psup.data(), npsup, &npsup));
  const size_t ncrs = psup.size();

  dh::TetMesh mesh(tet.data(), ntet, np);
  double *d_xyz = to_dev(xyz), *d_u = to_dev(u), *d_v = to_dev(v), *d_a = to_dev(a);
  double *d_R = to_dev(R), *d_b = to_dev(std::vector<double>(ndof, 0.0));

  bool all_same = true;
  std::printf("{");
  const char *kinds[4] = {"static", "newmark", "beuler", "stiffwarp"};
  for (int kind = 0; kind < 4; ++kind) {
    // (1) the reference's host matrix
    delfem2::CMatrixSparse<double> H;
    H.Initialize(np, 3, true);
    H.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
    H.setZero();
    std::vector<double> bH(ndof, 0.0);
    // (2) the same template on the device matrix (Merge found by ADL)
    dh::CMatrixSparse D;
    D.Initialize(np, 3, true);
    D.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
    D.setZero();
    std::vector<double> bD(ndof, 0.0);
    // (3) the device bulk assembly of the same kind
    dh::CMatrixSparse B;
    B.Initialize(np, 3, true);
    B.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
    B.setZero();
    dh::check(dfm2h_vec_set_zero(d_b, ndof));
    double t_ms = 0.0;
    bool okB = true;
    if (kind == 0) {
      delfem2::MergeLinSys_SolidLinear_Static_MeshTet3D(H, bH.data(), myu, lambda, rho, g,
                                                        xyz.data(), np, tet.data(), ntet,
                                                        u.data());
      const auto t0 = std::chrono::steady_clock::now();
      delfem2::MergeLinSys_SolidLinear_Static_MeshTet3D(D, bD.data(), myu, lambda, rho, g,
                                                        xyz.data(), np, tet.data(), ntet,
                                                        u.data());
      double *pv = nullptr;  // asking for the device arrays applies the staged blocks
      dh::check(dfm2h_bsr_info(D.h, nullptr, nullptr, &pv, nullptr));
      dh::check(dfm2h_synchronize());
      t_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                 .count();
      okB = dh::MergeLinSys_SolidLinear_Static_MeshTet3D(B, d_b, myu, lambda, rho, g, d_xyz,
                                                         mesh, d_u);
    } else if (kind == 1) {
      delfem2::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
          H, bH.data(), myu, lambda, rho, g, dt, gamma_nb, beta_nb, xyz.data(), np, tet.data(),
          ntet, u.data(), v.data(), a.data());
      delfem2::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
          D, bD.data(), myu, lambda, rho, g, dt, gamma_nb, beta_nb, xyz.data(), np, tet.data(),
          ntet, u.data(), v.data(), a.data());
      okB = dh::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
          B, d_b, myu, lambda, rho, g, dt, gamma_nb, beta_nb, d_xyz, mesh, d_u, d_v, d_a);
    } else if (kind == 2) {
      delfem2::MergeLinSys_SolidLinear_BEuler_MeshTet3D(H, bH.data(), myu, lambda, rho, g, dt,
                                                        xyz.data(), np, tet.data(), ntet,
                                                        u.data(), v.data());
      delfem2::MergeLinSys_SolidLinear_BEuler_MeshTet3D(D, bD.data(), myu, lambda, rho, g, dt,
                                                        xyz.data(), np, tet.data(), ntet,
                                                        u.data(), v.data());
      okB = dh::MergeLinSys_SolidLinear_BEuler_MeshTet3D(B, d_b, myu, lambda, rho, g, dt, d_xyz,
                                                         mesh, d_u, d_v);
    } else {
      delfem2::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(H, bH.data(), myu, lambda, rho, g,
                                                           dt, xyz.data(), np, tet.data(), ntet,
                                                           u.data(), v.data(), R);
      delfem2::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(D, bD.data(), myu, lambda, rho, g,
                                                           dt, xyz.data(), np, tet.data(), ntet,
                                                           u.data(), v.data(), R);
      okB = dh::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(B, d_b, myu, lambda, rho, g, dt,
                                                            d_xyz, mesh, d_u, d_v, d_R);
    }
    std::vector<double> vcD(9 * ncrs), vdD(9 * np), vcB(9 * ncrs), vdB(9 * np);
    D.download(vcD.data(), vdD.data());
    B.download(vcB.data(), vdB.data());
    const std::vector<double> bB = to_host(d_b, ndof);
    const bool template_same = same_bits(vcD, H.val_crs_) && same_bits(vdD, H.val_dia_) &&
                               same_bits(bD, bH);
    const bool bulk_same = okB && same_bits(vcB, H.val_crs_) && same_bits(vdB, H.val_dia_) &&
                           same_bits(bB, bH);
    all_same = all_same && template_same && bulk_same;
    std::printf("\"%s_template_same\": %s, \"%s_bulk_same\": %s, ", kinds[kind],
                template_same ? "true" : "false", kinds[kind], bulk_same ? "true" : "false");
    if (kind == 0) std::printf("\"static_template_ms\": %.3f, ", t_ms);
  }

  // a host array where a device pointer is required
  int refused = 0;
  {
    dh::CMatrixSparse M;
    M.Initialize(np, 3, true);
    M.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
    std::vector<int> flag(ndof, 0);
    try {
      M.SetFixedBC(flag.data());
    } catch (const dh::Error &e) {
      refused = e.code;
    }
  }
  std::printf("\"host_pointer_status\": %d, \"ntet\": %zu}\n", refused, ntet);
  for (void *p : {(void *)d_xyz, (void *)d_u, (void *)d_v, (void *)d_a, (void *)d_R,
                  (void *)d_b})
    dfm2h_free(p);
  return all_same && refused == DFM2H_ERR_INVALID ? 0 : 1;
}
