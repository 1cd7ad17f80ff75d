/*
 * delfem2_hip.hpp -- header-only C++ drop-in types over the C ABI
 * (dfm2_hip.h) for DelFEM2's implicit tet-FEM solve path.
 *
 * Namespace delfem2::hip mirrors the reference's names and signatures:
 *   CMatrixSparse        ls_block_sparse.h:24-204   (3x3 blocks, double)
 *   CPreconditionerILU   ls_ilu_block_sparse.h:21-60 (ILU(0), ILU(k))
 *   ViewAsVectorX<REAL>  view_vectorx.h:26-55 (a non-owning view of DEVICE memory)
 *   AddScaledVec / ScaleAndAddVec / AddMatVec / SolvePrecond / Dot
 *                        view_vectorx.h:61-109 (found by ADL)
 *   Merge<nrow,ncol,3,3,double>  ls_block_sparse.h:260-310 (found by ADL, so
 *                        the reference's MergeLinSys_* templates instantiate
 *                        on CMatrixSparse below)
 *   MergeLinSys_SolidLinear_Static_MeshTet3D  fem_solidlinear.h:192-237
 *   MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D fem_solidlinear.h:331-379
 *   MergeLinSys_SolidLinear_BEuler_MeshTet3D  fem_solidlinear.h:381-443
 *   MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D fem_solidlinear.h:445-530
 *   Solve_PBiCGStab      vecxitrsol.h:350-437 (device arrays)
 *   setRHS_Zero, XPlusAY vecxitrsol.cpp:226-296
 *   XPlusAYBZ, XPlusAYBZCW vecxitrsol.cpp:312-342
 *   Solve_PCG            lsitrsol.h:80-148, fused on the device
 *   Solve_CG             lsitrsol.h:36-74 (device views)
 *   LinearSystemSolver_BlockSparseILU  ls_solver_block_sparse_ilu.h:18-103
 *   MultiColorOrder      (no reference counterpart) node relabelling that
 *                        bounds the sweeps' levels by a colour count
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

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
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
    check(dfm2h_ilu_set_pattern_k(h, m.h, fill_level));
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

// ls_block_sparse.h:260-310 with the reference's signature.  An
// unqualified Merge<4, 4, 3, 3, double>(mat_A, aIP, aIP, emat, buffer) in
// the reference's element loops (fem_solidlinear.h:192-530) finds this one
// by ADL when mat_A is a delfem2::hip::CMatrixSparse: the blocks are staged
// and applied on the device in call order (bit-identical sums); the
// buffer is not needed.  False when a block is not in the pattern.
template <int nrow, int ncol, int ndimrow, int ndimcol, typename T>
bool Merge(CMatrixSparse &A, const unsigned int aIpRow[nrow], const unsigned int aIpCol[ncol],
           const T emat[nrow][ncol][ndimrow][ndimcol],
           std::vector<unsigned int> & /*merge_buffer*/) {
  static_assert(ndimrow == 3 && ndimcol == 3, "3x3 blocks only");
  static_assert(std::is_same<T, double>::value, "the device matrix holds doubles");
  const int st = dfm2h_bsr_merge(A.h, nrow, aIpRow, ncol, aIpCol, &emat[0][0][0][0]);
  if (st == DFM2H_ERR_PATTERN) return false;
  check(st);
  return true;
}

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

// Node relabelling for the level schedule (no reference counterpart; the C
// ABI's dfm2h_multicolor_order): new2old lists the nodes colour by colour,
// no two neighbours sharing a colour, from the symmetric one-ring pattern
// (psup_ind, psup) that JArray_PSuP_MeshElem builds.  Returns the colour
// count: the ILU(0) sweeps' levels in the relabelled mesh.
inline unsigned int MultiColorOrder(std::vector<unsigned int> &new2old,
                                    const std::vector<unsigned int> &psup_ind,
                                    const std::vector<unsigned int> &psup) {
  static const unsigned int kZero = 0;
  const size_t np = psup_ind.empty() ? 0 : psup_ind.size() - 1;
  new2old.resize(np);
  unsigned int nc = 0;
  check(dfm2h_multicolor_order(np, psup_ind.empty() ? &kZero : psup_ind.data(), psup.data(),
                               new2old.data(), &nc));
  return nc;
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

// fem_solidlinear.h:331-379 (the reference's argument order; device arrays;
// the mesh handle in place of nXYZ / aTet / nTet)
inline bool MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D(
    CMatrixSparse &mat_A, double *vec_b, const double myu, const double lambda, const double rho,
    const double g[3], const double dt_timestep, const double gamma_newmark,
    const double beta_newmark, const double *aXYZ, const TetMesh &mesh, const double *aVal,
    const double *aVelo, const double *aAcc, const double *myu_e = nullptr,
    const double *lambda_e = nullptr) {
  const int st = dfm2h_merge_linsys_solidlinear_newmark_tet(
      mat_A.h, vec_b, myu_e, lambda_e, myu, lambda, rho, g, dt_timestep, gamma_newmark,
      beta_newmark, aXYZ, mesh.h, aVal, aVelo, aAcc);
  if (st == DFM2H_ERR_PATTERN) return false;
  check(st);
  return true;
}

// fem_solidlinear.h:381-443
inline bool MergeLinSys_SolidLinear_BEuler_MeshTet3D(
    CMatrixSparse &mat_A, double *vec_b, const double myu, const double lambda, const double rho,
    const double *g, const double dt, const double *aXYZ, const TetMesh &mesh,
    const double *aDisp, const double *aVelo, const double *myu_e = nullptr,
    const double *lambda_e = nullptr) {
  const int st = dfm2h_merge_linsys_solidlinear_beuler_tet(mat_A.h, vec_b, myu_e, lambda_e, myu,
                                                           lambda, rho, g, dt, aXYZ, mesh.h,
                                                           aDisp, aVelo);
  if (st == DFM2H_ERR_PATTERN) return false;
  check(st);
  return true;
}

// fem_solidlinear.h:445-530 (aR: device rotations, 9 per node)
inline bool MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D(
    CMatrixSparse &mat_A, double *vec_b, const double myu, const double lambda, const double rho,
    const double g[3], const double dt, const double *aXYZ, const TetMesh &mesh,
    const double *aDisp, const double *aVelo, const double *aR, const double *myu_e = nullptr,
    const double *lambda_e = nullptr) {
  const int st = dfm2h_merge_linsys_solidstiffwarp_beuler_tet(
      mat_A.h, vec_b, myu_e, lambda_e, myu, lambda, rho, g, dt, aXYZ, mesh.h, aDisp, aVelo, aR);
  if (st == DFM2H_ERR_PATTERN) return false;
  check(st);
  return true;
}

// vecxitrsol.h:350-437 on device arrays (r_vec: b in, residual out)
inline std::vector<double> Solve_PBiCGStab(double *r_vec, double *x_vec, double conv_ratio_tol,
                                           unsigned int max_niter, const CMatrixSparse &mat,
                                           const CPreconditionerILU &ilu) {
  std::vector<double> hist(max_niter + 2);
  size_t nh = 0;
  check(dfm2h_solve_pbicgstab(mat.h, ilu.h, r_vec, x_vec, conv_ratio_tol, max_niter,
                              hist.data(), hist.size(), &nh));
  hist.resize(std::min(nh, hist.size()));
  return hist;
}

inline void XPlusAYBZ(double *X, size_t nDoF, const int *aBCFlag, double alpha, const double *Y,
                      double beta, const double *Z) {
  check(dfm2h_xplusaybz(X, nDoF, aBCFlag, alpha, Y, beta, Z));
}

inline void XPlusAYBZCW(double *X, size_t nDoF, const int *aBCFlag, double alpha, const double *Y,
                        double beta, const double *Z, double gamma, const double *W) {
  check(dfm2h_xplusaybzcw(X, nDoF, aBCFlag, alpha, Y, beta, Z, gamma, W));
}

// lsitrsol.h:36-74 Solve_CG (no preconditioner) on device views, the same
// recurrences and return value (relative residual per iteration); each Dot
// is a host synchronisation.
inline std::vector<double> Solve_CG(ViewAsVectorXd &&r_vec, ViewAsVectorXd &&u_vec,
                                    ViewAsVectorXd &&Ap_vec, ViewAsVectorXd &&p_vec,
                                    double conv_ratio_tol, unsigned int max_iteration,
                                    const CMatrixSparse &mat) {
  std::vector<double> conv;
  u_vec.setZero();
  double sq = Dot(r_vec, r_vec);
  if (sq < 1.0e-30) return conv;
  const double inv_sq0 = 1.0 / sq;
  p_vec = r_vec;
  for (unsigned int it = 0; it < max_iteration; ++it) {
    AddMatVec(Ap_vec, 0.0, 1.0, mat, p_vec);
    const double alpha = sq / Dot(p_vec, Ap_vec);
    AddScaledVec(u_vec, +alpha, p_vec);
    AddScaledVec(r_vec, -alpha, Ap_vec);
    const double sq_new = Dot(r_vec, r_vec);
    const double ratio = std::sqrt(sq_new * inv_sq0);
    conv.push_back(ratio);
    if (ratio < conv_ratio_tol) return conv;
    const double beta = sq_new / sq;
    sq = sq_new;
    ScaleAndAddVec(p_vec, beta, r_vec);
  }
  return conv;
}

// Device buffer owned by the facade below.
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() {
    if (p) dfm2h_free(p);
  }
  void resize(size_t count) {
    if (count == n) return;
    if (p) dfm2h_free(p);
    p = nullptr;
    n = count;
    if (n) {
      void *q = nullptr;
      check(dfm2h_malloc(&q, sizeof(T) * n));
      p = static_cast<T *>(q);
    }
  }
  void upload(const std::vector<T> &v) {
    resize(v.size());
    if (n) check(dfm2h_memcpy_htod(p, v.data(), sizeof(T) * n));
  }
  void download(std::vector<T> &v) const {
    v.resize(n);
    if (n) check(dfm2h_memcpy_dtoh(v.data(), p, sizeof(T) * n));
  }
  T *p = nullptr;
  size_t n = 0;
};

// ls_solver_block_sparse_ilu.h:18-103 LinearSystemSolver_BlockSparseILU, the
// caller-level facade.  Same members and calls; 3x3 blocks (ndim == 3).
// Element matrices are merged on the host as Merge (ls_block_sparse.h:
// 260-310) merges them -- the same additions in call order, into host copies
// of val_crs_ / val_dia_ -- and the solves upload those values and run on the
// device: SetFixedBC, CopyValue, Decompose and the fused Solve_PCG
// (Solve_PcgIlu: tol 1e-5, 1000 iterations), or Solve_CG (Solve_Cg: tol
// 1e-4, 300 iterations).  vec_r / vec_x / dof_bcflag stay host vectors as in
// the reference; Solve_PcgIlu also keeps its history in conv_hist (the
// reference prints it).
class LinearSystemSolver_BlockSparseILU {
 public:
  void Initialize(unsigned int nblk, unsigned int ndim, std::vector<unsigned int> &psup_ind,
                  std::vector<unsigned int> &psup) {
    if (ndim != 3) throw Error(DFM2H_ERR_INVALID, "only 3x3 blocks are supported");
    matrix.Initialize(nblk, ndim, true);
    matrix.SetPattern(psup_ind.data(), psup_ind.size(), psup.data(), psup.size());
    col_ind_ = psup_ind;
    row_ptr_ = psup;
    val_crs_.assign(row_ptr_.size() * 9, 0.0);
    val_dia_.assign(static_cast<size_t>(nblk) * 9, 0.0);
    dof_bcflag.assign(ndof(), 0);
    ilu_sparse.Initialize_ILUk(matrix, 0);
  }

  size_t nblk() const { return matrix.nrowblk_; }
  size_t ndim() const { return matrix.nrowdim_; }
  size_t ndof() const { return nblk() * ndim(); }

  void BeginMerge() {
    std::fill(val_crs_.begin(), val_crs_.end(), 0.0);
    std::fill(val_dia_.begin(), val_dia_.end(), 0.0);
    vec_r.assign(ndof(), 0.0);
  }

  // returns false when a block is missing from the pattern (the reference's
  // Merge result; its facade ignores it)
  template <int nrow, int ncol, int ndimrow, int ndimcol>
  bool Merge(const unsigned int *aIpRow, const unsigned int *aIpCol,
             const double emat[nrow][ncol][ndimrow][ndimcol]) {
    static_assert(ndimrow == 3 && ndimcol == 3, "3x3 blocks only");
    merge_buffer.resize(nblk(), ~0u);
    for (int ir = 0; ir < nrow; ++ir) {
      const unsigned int i = aIpRow[ir];
      for (unsigned int q = col_ind_[i]; q < col_ind_[i + 1]; ++q) merge_buffer[row_ptr_[q]] = q;
      bool ok = true;
      for (int jc = 0; jc < ncol && ok; ++jc) {
        const unsigned int j = aIpCol[jc];
        double *out;
        const double *in;
        if (i == j) {
          in = &emat[ir][ir][0][0];
          out = &val_dia_[static_cast<size_t>(i) * 9];
        } else {
          if (merge_buffer[j] == ~0u) {
            ok = false;
            break;
          }
          in = &emat[ir][jc][0][0];
          out = &val_crs_[static_cast<size_t>(merge_buffer[j]) * 9];
        }
        for (int k = 0; k < 9; ++k) out[k] += in[k];
      }
      for (unsigned int q = col_ind_[i]; q < col_ind_[i + 1]; ++q) merge_buffer[row_ptr_[q]] = ~0u;
      if (!ok) return false;
    }
    return true;
  }

  void AddValueToDiagonal(unsigned int iblk, unsigned int idim, double val) {
    val_dia_[static_cast<size_t>(iblk) * 9 + idim * 3 + idim] += val;
  }

  void Solve_Cg() {
    prepare();
    conv_hist = Solve_CG(ViewAsVectorXd(d_r.p, ndof()), ViewAsVectorXd(d_x.p, ndof()),
                         ViewAsVectorXd(d_t0.p, ndof()), ViewAsVectorXd(d_t1.p, ndof()), 1.0e-4,
                         300, matrix);
    finish();
  }

  void Solve_PcgIlu() {
    prepare();
    ilu_sparse.CopyValue(matrix);
    ilu_sparse.Decompose();
    conv_hist = Solve_PCG(ViewAsVectorXd(d_r.p, ndof()), ViewAsVectorXd(d_x.p, ndof()),
                          ViewAsVectorXd(d_t0.p, ndof()), ViewAsVectorXd(d_t1.p, ndof()), 1.0e-5,
                          1000, matrix, ilu_sparse);
    finish();
  }

 public:
  std::vector<double> vec_r;
  std::vector<double> vec_x;
  std::vector<int> dof_bcflag;
  std::vector<double> conv_hist;
  std::vector<unsigned int> merge_buffer;
  CMatrixSparse matrix;
  CPreconditionerILU ilu_sparse;
  // host copies of the pattern and of the merged values
  std::vector<unsigned int> col_ind_, row_ptr_;
  std::vector<double> val_crs_, val_dia_;

 private:
  // values and flags to the device, fixed BCs, rhs zeroed at fixed DOFs
  void prepare() {
    matrix.upload(val_crs_.data(), val_dia_.data());
    d_flag.upload(dof_bcflag);
    matrix.SetFixedBC(d_flag.p);
    for (size_t i = 0; i < vec_r.size(); ++i)
      if (dof_bcflag[i] != 0) vec_r[i] = 0.0;  // setRHS_Zero (vecxitrsol.cpp:286-296)
    d_r.upload(vec_r);
    d_x.resize(ndof());
    d_t0.resize(ndof());
    d_t1.resize(ndof());
  }
  void finish() {
    d_x.download(vec_x);
    d_r.download(vec_r);
  }
  DeviceArray<double> d_r, d_x, d_t0, d_t1;
  DeviceArray<int> d_flag;
};

}  // namespace hip

// Makes the name Merge a template visible in namespace delfem2, so that the
// reference's element loops (which call Merge<...>(mat_A, ...) unqualified)
// parse under C++17 even where ls_block_sparse.h is not included before
// them; the call itself resolves by overload (delfem2::hip::CMatrixSparse).
using hip::Merge;

}  // namespace delfem2

#endif  // DELFEM2_HIP_HPP
