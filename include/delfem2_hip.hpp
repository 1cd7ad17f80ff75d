/*
 * delfem2_hip.hpp -- header-only C++ drop-in types over the C ABI
 * (dfm2_hip.h) for DelFEM2's implicit tet-FEM solve path.
 *
 * Namespace delfem2::hip mirrors the reference's names and signatures:
 *   CMatrixSparse        ls_block_sparse.h:24-204   (3x3 blocks, double)
 *   CPreconditionerILU   ls_ilu_block_sparse.h:21-60 (ILU(0))
 *   ViewAsVectorX<REAL>  view_vectorx.h:26-55 (a non-owning view of DEVICE memory)
 *   AddScaledVec / ScaleAndAddVec / AddMatVec / SolvePrecond / Dot
 *                        view_vectorx.h:61-109 (found by ADL)
 *   MergeLinSys_SolidLinear_Static_MeshTet3D  fem_solidlinear.h:192-237
 *   setRHS_Zero, XPlusAY vecxitrsol.cpp:226-296
 *   Solve_PCG            lsitrsol.h:80-148, fused on the device
 *
 * With these types the reference's own template
 *   delfem2::Solve_PCG(VEC&&, VEC&&, VEC&&, VEC&&, double, unsigned, const MAT&, const PREC&)
 * instantiates unchanged (each Dot returns a host double: three host
 * synchronisations per iteration); an unqualified call, found by ADL,
 * selects delfem2::hip::Solve_PCG below, which runs the whole iteration on
 * the device.  Failures other than the reference's bool returns throw
 * delfem2::hip::Error.
 */
#ifndef DELFEM2_HIP_HPP
#define DELFEM2_HIP_HPP

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "dfm2_hip.h"

namespace delfem2 {
namespace hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int status) {
  if (status != DFM2H_OK) throw Error(status, dfm2h_last_error());
}

class CMatrixSparse {
 public:
  CMatrixSparse() noexcept = default;
  CMatrixSparse(const CMatrixSparse &) = delete;
  CMatrixSparse &operator=(const CMatrixSparse &) = delete;
  ~CMatrixSparse() { Clear(); }

  void Clear() {
    if (h) dfm2h_bsr_destroy(h);
    h = nullptr;
    nrowblk_ = ncolblk_ = nrowdim_ = ncoldim_ = 0;
  }
  void Initialize(size_t nblk, unsigned int len, bool is_dia) {
    Clear();
    check(dfm2h_bsr_create(&h, nblk, len, is_dia ? 1 : 0));
    nrowblk_ = ncolblk_ = static_cast<unsigned int>(nblk);
    nrowdim_ = ncoldim_ = len;
  }
  // host arrays, copied (ls_block_sparse.h:72-85)
  void SetPattern(const unsigned int *colind, size_t ncolind, const unsigned int *rowptr,
                  size_t nrowptr) {
    check(dfm2h_bsr_set_pattern(h, colind, ncolind, rowptr, nrowptr));
  }
  bool setZero() {
    check(dfm2h_bsr_set_zero(h));
    return true;
  }
  // device pointers: y = alpha*A*x + beta*y
  void MatVec(double *y, double alpha, const double *x, double beta) const {
    check(dfm2h_bsr_matvec(h, y, alpha, x, beta));
  }
  // device array of 3*nblk flags
  void SetFixedBC(const int *bc_flag) { check(dfm2h_bsr_set_fixed_bc(h, bc_flag)); }
  // val_crs_ / val_dia_ live in HBM: explicit host copies
  void upload(const double *val_crs, const double *val_dia) {
    check(dfm2h_bsr_upload(h, val_crs, val_dia));
  }
  void download(double *val_crs, double *val_dia) const {
    check(dfm2h_bsr_download(h, val_crs, val_dia));
  }
  size_t ncrs() const {
    size_t n = 0;
    check(dfm2h_bsr_info(h, nullptr, &n, nullptr, nullptr));
    return n;
  }

  dfm2h_bsr h = nullptr;
  unsigned int nrowblk_ = 0, ncolblk_ = 0, nrowdim_ = 0, ncoldim_ = 0;
};

class CPreconditionerILU {
 public:
  CPreconditionerILU() { check(dfm2h_ilu_create(&h)); }
  CPreconditionerILU(const CPreconditionerILU &) = delete;
  CPreconditionerILU &operator=(const CPreconditionerILU &) = delete;
  ~CPreconditionerILU() {
    if (h) dfm2h_ilu_destroy(h);
  }
  void SetPattern0(const CMatrixSparse &m) { check(dfm2h_ilu_set_pattern0(h, m.h)); }
  void Initialize_ILUk(const CMatrixSparse &m, int fill_level) {
    if (fill_level != 0) throw Error(DFM2H_ERR_INVALID, "only ILU(0) is implemented");
    SetPattern0(m);
  }
  void CopyValue(const CMatrixSparse &m) { check(dfm2h_ilu_copy_value(h, m.h)); }
  bool Decompose() {
    int ok = 0;
    check(dfm2h_ilu_decompose(h, &ok, &n_singular));
    return ok != 0;
  }
  void SolvePrecond(double *vec) const { check(dfm2h_ilu_solve_precond(h, vec)); }
  void ForwardSubstitution(double *vec) const { check(dfm2h_ilu_forward(h, vec)); }
  void BackwardSubstitution(double *vec) const { check(dfm2h_ilu_backward(h, vec)); }

  dfm2h_ilu h = nullptr;
  int n_singular = 0;
};

// Non-owning view of a device vector (view_vectorx.h:26-55).
template <typename REAL>
class ViewAsVectorX {
 public:
  ViewAsVectorX(REAL *p_, std::size_t n_) : p(p_), n(n_) {}
  REAL dot(const ViewAsVectorX &rhs) const {
    double d = 0.0;
    check(dfm2h_vec_dot(p, rhs.p, n, &d));
    return d;
  }
  void operator=(const ViewAsVectorX &rhs) { check(dfm2h_vec_copy(p, rhs.p, n)); }
  void setZero() { check(dfm2h_vec_set_zero(p, n)); }

  REAL *const p;
  const std::size_t n;
};
using ViewAsVectorXd = ViewAsVectorX<double>;

inline void AddScaledVec(ViewAsVectorXd &y, double alpha, const ViewAsVectorXd &x) {
  check(dfm2h_vec_add_scaled(y.p, alpha, x.p, y.n));
}
inline void ScaleAndAddVec(ViewAsVectorXd &y, double beta, const ViewAsVectorXd &x) {
  check(dfm2h_vec_scale_and_add(y.p, beta, x.p, y.n));
}
inline void AddMatVec(ViewAsVectorXd &lhs, double scale_lhs, double scale_rhs,
                      const CMatrixSparse &mat, const ViewAsVectorXd &rhs) {
  mat.MatVec(lhs.p, scale_rhs, rhs.p, scale_lhs);
}
inline void SolvePrecond(ViewAsVectorXd &Pr_vec, const CPreconditionerILU &ilu) {
  ilu.SolvePrecond(Pr_vec.p);
}
inline double Dot(const ViewAsVectorXd &x, const ViewAsVectorXd &y) { return x.dot(y); }

// Fused device Solve_PCG (lsitrsol.h:80-148): same arguments and return
// value; Pr and p are scratch in the reference and are not touched here.
inline std::vector<double> Solve_PCG(ViewAsVectorXd &&r_vec, ViewAsVectorXd &&x_vec,
                                     ViewAsVectorXd && /*Pr_vec*/, ViewAsVectorXd && /*p_vec*/,
                                     double conv_ratio_tol, unsigned int max_nitr,
                                     const CMatrixSparse &mat, const CPreconditionerILU &ilu) {
  std::vector<double> hist(static_cast<size_t>(max_nitr) + 2);
  size_t nh = 0;
  check(dfm2h_solve_pcg(mat.h, ilu.h, r_vec.p, x_vec.p, conv_ratio_tol, max_nitr, hist.data(),
                        hist.size(), &nh));
  hist.resize(nh);
  return hist;
}

// fem_solidlinear.h:192-237 on device arrays; per-element Lame arrays may be
// null (then the scalars apply).  Returns false when an element edge has no
// block in the pattern (the reference's Merge result).
class TetMesh {
 public:
  TetMesh(const unsigned int *tet_host, size_t ntet, size_t np) {
    check(dfm2h_tetmesh_create(&h, tet_host, ntet, np));
  }
  TetMesh(const TetMesh &) = delete;
  TetMesh &operator=(const TetMesh &) = delete;
  ~TetMesh() {
    if (h) dfm2h_tetmesh_destroy(h);
  }
  dfm2h_tetmesh h = nullptr;
};

inline bool MergeLinSys_SolidLinear_Static_MeshTet3D(
    CMatrixSparse &mat_A, double *vec_b, const double myu, const double lambda, const double rho,
    const double *g, const double *aXYZ, const TetMesh &mesh, const double *aDisp,
    const double *myu_e = nullptr, const double *lambda_e = nullptr) {
  const int st = dfm2h_merge_linsys_solidlinear_static_tet(mat_A.h, vec_b, myu_e, lambda_e, myu,
                                                           lambda, rho, g, aXYZ, mesh.h, aDisp);
  if (st == DFM2H_ERR_PATTERN) return false;
  check(st);
  return true;
}

inline void setRHS_Zero(double *vec_b, const int *aBCFlag, size_t n, int iflag_nonzero) {
  check(dfm2h_set_rhs_zero(vec_b, aBCFlag, n, iflag_nonzero));
}

inline void XPlusAY(double *X, size_t nDoF, const int *aBCFlag, double alpha, const double *Y) {
  check(dfm2h_xplusay(X, nDoF, aBCFlag, alpha, Y));
}

}  // namespace hip
}  // namespace delfem2

#endif  // DELFEM2_HIP_HPP
