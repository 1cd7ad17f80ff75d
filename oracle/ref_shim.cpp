// ref_shim.cpp -- extern "C" entry points over the REFERENCE's own headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (header-only mode) from the reference
// sources where they lie, by oracle/Makefile, into oracle/_ref/libdfm2ref.so,
// which is git-ignored.  It is used to pin the oracle restatement
// (oracle.c) bit for bit, to write the golden fixtures under tests/golden/,
// and as the CPU baseline.  Nothing in the product links it.
//
// Each function drives one reference routine; the handle keeps the
// reference's CMatrixSparse<double> and CPreconditionerILU<double> objects.

#include <cstring>
#include <vector>

// ls_block_sparse.h first: fem_solidlinear.h's drivers call Merge<> and only
// include it themselves in static-library mode.
#include "delfem2/ls_block_sparse.h"
#include "delfem2/ls_ilu_block_sparse.h"
#include "delfem2/fem_solidlinear.h"
#include "delfem2/femutil.h"
#include "delfem2/gridvoxel.h"
#include "delfem2/jagarray.h"
#include "delfem2/lsitrsol.h"
#include "delfem2/msh_topology_uniform.h"
#include "delfem2/vecxitrsol.h"
#include "delfem2/view_vectorx.h"
#include "delfem2/svd3.h"
#include "delfem2/mat3.h"
#include "delfem2/mat3_funcs.h"
#include "delfem2/vec3.h"

namespace dfm2 = delfem2;

struct RefSys {
  dfm2::CMatrixSparse<double> A;
  dfm2::CPreconditionerILU<double> ilu;
  std::vector<unsigned int> buf;
};

extern "C" {

// gridvoxel.cpp:194-271 (all voxels on); integer coordinates.
void ref_mesh_tet_voxel(unsigned nx, unsigned ny, unsigned nz, double *xyz, unsigned *tet) {
  std::vector<double> aXYZ;
  std::vector<int> aTet;
  dfm2::MeshTet3D_VoxelGrid(aXYZ, aTet, nx, ny, nz, std::vector<int>(nx * ny * nz, 1));
  std::memcpy(xyz, aXYZ.data(), sizeof(double) * aXYZ.size());
  for (size_t i = 0; i < aTet.size(); ++i) tet[i] = static_cast<unsigned>(aTet[i]);
}

// msh_topology_uniform.cpp:301-316 + jagarray.cpp:72-92
size_t ref_psup(const unsigned *tet, size_t ntet, size_t np, unsigned *psup_ind, unsigned *psup) {
  std::vector<unsigned int> ind, arr;
  dfm2::JArray_PSuP_MeshElem(ind, arr, tet, ntet, 4, np);
  dfm2::JArray_Sort(ind, arr);
  std::memcpy(psup_ind, ind.data(), sizeof(unsigned) * ind.size());
  if (psup) std::memcpy(psup, arr.data(), sizeof(unsigned) * arr.size());
  return arr.size();
}

void *ref_create(unsigned nblk, const unsigned *colind, const unsigned *rowptr) {
  auto *s = new RefSys;
  s->A.Initialize(nblk, 3, true);
  s->A.SetPattern(colind, nblk + 1, rowptr, colind[nblk]);
  s->ilu.SetPattern0(s->A);
  s->A.setZero();
  return s;
}
void ref_destroy(void *h) { delete static_cast<RefSys *>(h); }

void ref_get_matrix(void *h, double *vcrs, double *vdia) {
  auto *s = static_cast<RefSys *>(h);
  std::memcpy(vcrs, s->A.val_crs_.data(), sizeof(double) * s->A.val_crs_.size());
  std::memcpy(vdia, s->A.val_dia_.data(), sizeof(double) * s->A.val_dia_.size());
}
void ref_set_matrix(void *h, const double *vcrs, const double *vdia) {
  auto *s = static_cast<RefSys *>(h);
  std::memcpy(s->A.val_crs_.data(), vcrs, sizeof(double) * s->A.val_crs_.size());
  std::memcpy(s->A.val_dia_.data(), vdia, sizeof(double) * s->A.val_dia_.size());
}

// fem_solidlinear.h:192-237 with a uniform (myu, lambda): the reference driver itself.
void ref_assemble_uniform(void *h, double *vec_b, double myu, double lambda, double rho,
                          const double *g, const double *xyz, size_t np, const unsigned *tet,
                          size_t ntet, const double *disp) {
  auto *s = static_cast<RefSys *>(h);
  dfm2::MergeLinSys_SolidLinear_Static_MeshTet3D(s->A, vec_b, myu, lambda, rho, g, xyz, np, tet,
                                                  ntet, disp);
}

// The same element loop with per-element (myu, lambda): FetchData, TetVolume3D,
// EMat_SolidLinear_Static_Tet and Merge<4,4,3,3> called in ascending element order.
int ref_assemble_per_element(void *h, double *vec_b, const double *myu_e,
                             const double *lambda_e, double rho, const double *g,
                             const double *xyz, const unsigned *tet, size_t ntet,
                             const double *disp) {
  auto *s = static_cast<RefSys *>(h);
  s->buf.assign(s->A.nrowblk_, UINT_MAX);
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *aIP = tet + iel * 4;
    double P[4][3];
    dfm2::FetchData<4, 3>(P, aIP, xyz);
    double disps[4][3];
    dfm2::FetchData<4, 3>(disps, aIP, disp);
    double eres[4][3];
    const double vol = dfm2::femutil::TetVolume3D(P[0], P[1], P[2], P[3]);
    for (auto &e : eres) {
      e[0] = vol * rho * g[0] * 0.25;
      e[1] = vol * rho * g[1] * 0.25;
      e[2] = vol * rho * g[2] * 0.25;
    }
    double emat[4][4][3][3];
    for (int i = 0; i < 144; ++i) (&emat[0][0][0][0])[i] = 0.0;
    dfm2::EMat_SolidLinear_Static_Tet(emat, eres, myu_e[iel], lambda_e[iel], P, disps, true);
    for (int ino = 0; ino < 4; ++ino) {
      const unsigned ip = aIP[ino];
      vec_b[ip * 3 + 0] += eres[ino][0];
      vec_b[ip * 3 + 1] += eres[ino][1];
      vec_b[ip * 3 + 2] += eres[ino][2];
    }
    if (!dfm2::Merge<4, 4, 3, 3, double>(s->A, aIP, aIP, emat, s->buf)) status = 1;
  }
  return status;
}

void ref_emat_tet(double *emat, double *eres, double myu, double lambda, const double *P,
                  const double *disp) {
  dfm2::EMat_SolidLinear_Static_Tet(*reinterpret_cast<double(*)[4][4][3][3]>(emat),
                                    *reinterpret_cast<double(*)[4][3]>(eres), myu, lambda,
                                    *reinterpret_cast<const double(*)[4][3]>(P),
                                    *reinterpret_cast<const double(*)[4][3]>(disp), true);
}

void ref_tet_dldx(double *dldx, const double *p0, const double *p1, const double *p2,
                  const double *p3) {
  double a[4];
  dfm2::TetDlDx(reinterpret_cast<double(*)[3]>(dldx), a, p0, p1, p2, p3);
}

void ref_set_fixed_bc(void *h, const int *flag) {
  static_cast<RefSys *>(h)->A.SetFixedBC(flag);
}

void ref_set_rhs_zero(double *b, const int *flag, size_t n, int iflag_nonzero) {
  std::vector<double> vb(b, b + n);
  std::vector<int> vf(flag, flag + n);
  dfm2::setRHS_Zero(vb, vf, iflag_nonzero);
  std::memcpy(b, vb.data(), sizeof(double) * n);
}

void ref_matvec(void *h, double *y, double alpha, const double *x, double beta) {
  static_cast<RefSys *>(h)->A.MatVec(y, alpha, x, beta);
}

// ls_ilu_block_sparse.cpp:849-999 (lev_fill 0: SetPattern0)
size_t ref_ilu_init_k(void *h, int lev_fill, unsigned *colind) {
  auto *s = static_cast<RefSys *>(h);
  s->ilu.Initialize_ILUk(s->A, lev_fill);
  if (colind) std::memcpy(colind, s->ilu.colInd.data(), sizeof(unsigned) * s->ilu.colInd.size());
  return s->ilu.rowPtr.size();
}

// ls_ilu_block_sparse.cpp:1009-1047 then :152-420
int ref_ilu_factor(void *h) {
  auto *s = static_cast<RefSys *>(h);
  s->ilu.CopyValue(s->A);
  return s->ilu.Decompose() ? 1 : 0;
}

void ref_ilu_get(void *h, unsigned *rowptr, unsigned *diaind, double *vcrs, double *vdia) {
  auto *s = static_cast<RefSys *>(h);
  if (rowptr) std::memcpy(rowptr, s->ilu.rowPtr.data(), sizeof(unsigned) * s->ilu.rowPtr.size());
  if (diaind) std::memcpy(diaind, s->ilu.m_diaInd.data(), sizeof(unsigned) * s->ilu.m_diaInd.size());
  if (vcrs) std::memcpy(vcrs, s->ilu.valCrs.data(), sizeof(double) * s->ilu.valCrs.size());
  if (vdia) std::memcpy(vdia, s->ilu.valDia.data(), sizeof(double) * s->ilu.valDia.size());
}

void ref_ilu_forward(void *h, double *vec) { static_cast<RefSys *>(h)->ilu.ForwardSubstitution(vec); }
void ref_ilu_backward(void *h, double *vec) { static_cast<RefSys *>(h)->ilu.BackwardSubstitution(vec); }
void ref_ilu_solve_precond(void *h, double *vec) { static_cast<RefSys *>(h)->ilu.SolvePrecond(vec); }

// lsitrsol.h:80-148 through the ViewAsVectorXd operator set.
size_t ref_solve_pcg(void *h, double *r, double *x, double tol, unsigned max_nitr, double *hist,
                     size_t hist_cap) {
  auto *s = static_cast<RefSys *>(h);
  const size_t n = static_cast<size_t>(s->A.nrowblk_) * 3;
  std::vector<double> t0(n), t1(n);
  std::vector<double> hv = dfm2::Solve_PCG(
      dfm2::ViewAsVectorXd(r, n), dfm2::ViewAsVectorXd(x, n), dfm2::ViewAsVectorXd(t0.data(), n),
      dfm2::ViewAsVectorXd(t1.data(), n), tol, max_nitr, s->A, s->ilu);
  for (size_t i = 0; i < hv.size() && i < hist_cap; ++i) hist[i] = hv[i];
  return hv.size();
}

void ref_xplusay(double *X, size_t n, const int *flag, double alpha, const double *Y) {
  std::vector<double> vx(X, X + n), vy(Y, Y + n);
  std::vector<int> vf(flag, flag + n);
  dfm2::XPlusAY(vx, n, vf, alpha, vy);
  std::memcpy(X, vx.data(), sizeof(double) * n);
}

// fem_solidlinear.h:331-379: the reference's Newmark-beta driver itself
// (uniform myu, lambda).
void ref_assemble_newmark(void *h, double *vec_b, double myu, double lambda, double rho,
                          const double *g, double dt, double gamma, double beta,
                          const double *xyz, size_t np, const unsigned *tet, size_t ntet,
                          const double *disp, const double *velo, const double *acc) {
  auto *s = static_cast<RefSys *>(h);
  const double gg[3] = {g[0], g[1], g[2]};
  dfm2::MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(s->A, vec_b, myu, lambda, rho, gg, dt,
                                                      gamma, beta, xyz, np, tet, ntet, disp,
                                                      velo, acc);
}

// The same element loop with per-element (myu, lambda): FetchData,
// EMat_SolidLinear_NewmarkBeta_MeshTet3D (fem_solidlinear.cpp:437-490,
// is_initial_iter = true) and Merge<4,4,3,3> in ascending element order.
int ref_assemble_newmark_per_element(void *h, double *vec_b, const double *myu_e,
                                     const double *lambda_e, double rho, const double *g,
                                     double dt, double gamma, double beta, const double *xyz,
                                     const unsigned *tet, size_t ntet, const double *disp,
                                     const double *velo, const double *acc) {
  auto *s = static_cast<RefSys *>(h);
  s->buf.assign(s->A.nrowblk_, UINT_MAX);
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *aIP = tet + iel * 4;
    double P[4][3], U[4][3], V[4][3], Ac[4][3];
    dfm2::FetchData<4, 3>(P, aIP, xyz);
    dfm2::FetchData<4, 3>(U, aIP, disp);
    dfm2::FetchData<4, 3>(V, aIP, velo);
    dfm2::FetchData<4, 3>(Ac, aIP, acc);
    double eres[4][3], emat[4][4][3][3];
    dfm2::EMat_SolidLinear_NewmarkBeta_MeshTet3D(eres, emat, myu_e[iel], lambda_e[iel], rho,
                                                 g[0], g[1], g[2], dt, gamma, beta, U, V, Ac,
                                                 P, true);
    for (int ino = 0; ino < 4; ++ino) {
      const unsigned ip = aIP[ino];
      vec_b[ip * 3 + 0] += eres[ino][0];
      vec_b[ip * 3 + 1] += eres[ino][1];
      vec_b[ip * 3 + 2] += eres[ino][2];
    }
    if (!dfm2::Merge<4, 4, 3, 3, double>(s->A, aIP, aIP, emat, s->buf)) status = 1;
  }
  return status;
}

// fem_solidlinear.h:381-443: the reference's backward-Euler driver itself.
void ref_assemble_beuler(void *h, double *vec_b, double myu, double lambda, double rho,
                         const double *g, double dt, const double *xyz, size_t np,
                         const unsigned *tet, size_t ntet, const double *disp,
                         const double *velo) {
  auto *s = static_cast<RefSys *>(h);
  dfm2::MergeLinSys_SolidLinear_BEuler_MeshTet3D(s->A, vec_b, myu, lambda, rho, g, dt, xyz,
                                                 static_cast<unsigned>(np), tet,
                                                 static_cast<unsigned>(ntet), disp, velo);
}

// vecxitrsol.cpp:312-325 and :327-342
void ref_xplusaybz(double *X, size_t n, const int *flag, double alpha, const double *Y,
                   double beta, const double *Z) {
  std::vector<double> vx(X, X + n), vy(Y, Y + n), vz(Z, Z + n);
  std::vector<int> vf(flag, flag + n);
  dfm2::XPlusAYBZ(vx, n, vf, alpha, vy, beta, vz);
  std::memcpy(X, vx.data(), sizeof(double) * n);
}
void ref_xplusaybzcw(double *X, size_t n, const int *flag, double alpha, const double *Y,
                     double beta, const double *Z, double gamma, const double *W) {
  std::vector<double> vx(X, X + n), vy(Y, Y + n), vz(Z, Z + n), vw(W, W + n);
  std::vector<int> vf(flag, flag + n);
  dfm2::XPlusAYBZCW(vx, n, vf, alpha, vy, beta, vz, gamma, vw);
  std::memcpy(X, vx.data(), sizeof(double) * n);
}

// svd3.cpp:291-305
void ref_rot_polar_decomp(double *R, const double *m, int nitr) {
  dfm2::GetRotPolarDecomp(R, m, nitr);
}

// 603_Fem3DStiffwarp/main.cpp:80-106 (an example-file function: the same
// steps written here over the reference's CVec3d / CMat3d /
// Mat3_OuterProduct / GetRotPolarDecomp)
void ref_rotation_at_points(double *R, const double *xyz, const double *disp,
                            const unsigned *psup_ind, const unsigned *psup, size_t np) {
  for (unsigned int ip = 0; ip < np; ++ip) {
    dfm2::CVec3d Pi(xyz[ip * 3 + 0], xyz[ip * 3 + 1], xyz[ip * 3 + 2]);
    dfm2::CVec3d pi(xyz[ip * 3 + 0] + disp[ip * 3 + 0], xyz[ip * 3 + 1] + disp[ip * 3 + 1],
                    xyz[ip * 3 + 2] + disp[ip * 3 + 2]);
    dfm2::CMat3d A;
    A.setZero();
    for (unsigned int jjp = psup_ind[ip]; jjp < psup_ind[ip + 1]; ++jjp) {
      const unsigned int jp = psup[jjp];
      dfm2::CVec3d Pj(xyz[jp * 3 + 0], xyz[jp * 3 + 1], xyz[jp * 3 + 2]);
      dfm2::CVec3d pj(xyz[jp * 3 + 0] + disp[jp * 3 + 0], xyz[jp * 3 + 1] + disp[jp * 3 + 1],
                      xyz[jp * 3 + 2] + disp[jp * 3 + 2]);
      A += dfm2::Mat3_OuterProduct(pj - pi, Pj - Pi);
    }
    dfm2::GetRotPolarDecomp(R + ip * 9, A.p_, 100);
  }
}

// fem_solidlinear.h:445-530: the reference's stiffness-warping driver itself.
void ref_assemble_stiffwarp(void *h, double *vec_b, double myu, double lambda, double rho,
                            const double *g, double dt, const double *xyz, size_t np,
                            const unsigned *tet, size_t ntet, const double *disp,
                            const double *velo, const double *R) {
  auto *s = static_cast<RefSys *>(h);
  const std::vector<double> aR(R, R + np * 9);
  const double gg[3] = {g[0], g[1], g[2]};
  dfm2::MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(s->A, vec_b, myu, lambda, rho, gg, dt, xyz,
                                                    static_cast<unsigned>(np), tet,
                                                    static_cast<unsigned>(ntet), disp, velo, aR);
}

// vecxitrsol.h:350-437
size_t ref_solve_pbicgstab(void *h, double *r, double *x, double tol, unsigned max_niter,
                           double *hist, size_t hist_cap) {
  auto *s = static_cast<RefSys *>(h);
  std::vector<double> hv = dfm2::Solve_PBiCGStab(r, x, tol, max_niter, s->A, s->ilu);
  for (size_t i = 0; i < hv.size() && i < hist_cap; ++i) hist[i] = hv[i];
  return hv.size();
}

}  // extern "C"
