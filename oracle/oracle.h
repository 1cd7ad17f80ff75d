/*
 * oracle.h -- CPU restatement of DelFEM2's implicit tet-FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (delfem2_amd/) never calls it.
 *
 * Every function restates one reference routine with the same floating-point
 * association (compile with -ffp-contract=off), so that its outputs are
 * bit-identical to the reference on the same inputs.  Citations are relative
 * to the reference tree's include/delfem2/.
 *
 * Layout (identical to CMatrixSparse<double> with nrowdim=ncoldim=3):
 *   colind[nblk+1]   row-start offsets       (reference member col_ind_)
 *   rowptr[ncrs]     block column indices    (reference member row_ptr_)
 *   vcrs[9*ncrs]     off-diagonal 3x3 blocks, row-major
 *   vdia[9*nblk]     diagonal 3x3 blocks, row-major
 *   vectors          AoS x[3*node+d]
 */
#ifndef DFM2_ORACLE_H
#define DFM2_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gridvoxel.cpp:194-271 MeshTet3D_VoxelGrid, all voxels filled.
 * xyz: 3*(nx+1)*(ny+1)*(nz+1) integer-valued coordinates; tet: 4*6*nx*ny*nz. */
void orc_mesh_tet_voxel(unsigned nx, unsigned ny, unsigned nz,
                        double *xyz, unsigned *tet);

/* msh_topology_uniform.cpp:301-316 JArray_PSuP_MeshElem + jagarray.cpp:72-92
 * JArray_Sort.  psup_ind[np+1] is always filled; psup (nullable) receives the
 * sorted one-ring (self excluded).  Returns the number of entries. */
size_t orc_psup(const unsigned *elem, size_t nelem, unsigned nnoel, size_t np,
                unsigned *psup_ind, unsigned *psup);

/* femutil.cpp:202-212 TetVolume3D and femutil.cpp:215-258 TetDlDx. */
double orc_tet_volume(const double *p0, const double *p1, const double *p2,
                      const double *p3);
void orc_tet_dldx(double dldx[4][3], const double *p0, const double *p1,
                  const double *p2, const double *p3);

/* fem_solidlinear.cpp:409-435 EMat_SolidLinear_Static_Tet (is_add=true on a
 * zeroed emat) with AddEmat_LinearSolid3<4> (fem_solidlinear.cpp:55-79) and
 * AddEmatEvecScale3<4> (femutil.h:252-268).  P, disp: [4][3]. */
void orc_emat_tet(double emat[4][4][3][3], double eres[4][3], double myu,
                  double lambda, const double P[4][3], const double disp[4][3]);

/* fem_solidlinear.h:192-237 MergeLinSys_SolidLinear_Static_MeshTet3D with a
 * per-element (myu_e, lambda_e) pair, merged through Merge<4,4,3,3>
 * (ls_block_sparse.h:260-310) in ascending element order.  Additive into
 * vcrs/vdia/vec_b.  disp may be NULL (treated as zeros).
 * Returns 0, or 1 when a block is missing from the pattern. */
int orc_assemble_tet_static(unsigned nblk, const unsigned *colind,
                            const unsigned *rowptr, double *vcrs, double *vdia,
                            double *vec_b, const double *myu_e,
                            const double *lambda_e, double rho, const double *g,
                            const double *xyz, const unsigned *tet, size_t ntet,
                            const double *disp);

/* ls_block_sparse.h:147-151 SetFixedBC -> ls_block_sparse.cpp:397-472. */
void orc_set_fixed_bc(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, double *vcrs, double *vdia,
                      const int *bc_flag);

/* vecxitrsol.cpp:286-296 setRHS_Zero. */
void orc_set_rhs_zero(double *b, const int *flag, size_t n, int iflag_nonzero);

/* ls_block_sparse.cpp:215-273 MatVec -> :103-133 MatVec_MatSparseCRS_Blk33:
 * y = alpha*A*x + beta*y. */
void orc_matvec(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                const double *vcrs, const double *vdia, double *y,
                double alpha, const double *x, double beta);

/* ls_ilu_block_sparse.cpp:1057-1090 SetPattern0: rowptr_out gets the sorted
 * copy of rowptr, diaind[i] the first slot whose column is > i. */
void orc_ilu_set_pattern0(unsigned nblk, const unsigned *colind,
                          const unsigned *rowptr, unsigned *rowptr_out,
                          unsigned *diaind);

/* ls_ilu_block_sparse.cpp:1009-1047 CopyValue. */
size_t orc_ilu_pattern_k(unsigned nblk, const unsigned *a_colind, const unsigned *a_rowptr,
                         int lev_fill, unsigned *colind, unsigned *rowptr, unsigned *diaind,
                         size_t cap);
void orc_ilu_copy_value(unsigned nblk, const unsigned *ilu_colind,
                        const unsigned *ilu_rowptr, double *ilu_vcrs,
                        double *ilu_vdia, const unsigned *a_colind,
                        const unsigned *a_rowptr, const double *a_vcrs,
                        const double *a_vdia);

/* ls_ilu_block_sparse.cpp:152-420 Decompose() (3x3 branch 286-356, with
 * CalcInvMat3 :84-103).  Returns 1 (true) or 0 (false: more than 10
 * singular pivots; the factorisation stops there, as in the reference).
 * *n_singular receives the number of singular pivots met. */
int orc_ilu_decompose(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, const unsigned *diaind,
                      double *vcrs, double *vdia, int *n_singular);

/* ls_ilu_block_sparse.cpp:544-573 / :721-749 (3x3 branches). In place. */
void orc_ilu_forward(unsigned nblk, const unsigned *colind,
                     const unsigned *rowptr, const unsigned *diaind,
                     const double *vcrs, const double *vdia, double *vec);
void orc_ilu_backward(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, const unsigned *diaind,
                      const double *vcrs, double *vec);

/* view_vectorx.h:32-39 dot (sequential sum from 0). */
double orc_dot(const double *x, const double *y, size_t n);

/* lsitrsol.h:80-148 Solve_PCG on (A, ILU) with the view_vectorx.h:61-109
 * operator set.  r is overwritten with the final residual, x receives the
 * solution.  hist receives the residual-norm history (absolute norms; entry
 * 0 is |b|); returns its length (at most hist_cap is written). */
size_t orc_solve_pcg(double *r, double *x, double tol, unsigned max_nitr,
                     unsigned nblk, const unsigned *a_colind,
                     const unsigned *a_rowptr, const double *a_vcrs,
                     const double *a_vdia, const unsigned *l_colind,
                     const unsigned *l_rowptr, const unsigned *l_diaind,
                     const double *l_vcrs, const double *l_vdia, double *hist,
                     size_t hist_cap);

/* vecxitrsol.cpp:226-237 XPlusAY: X[i] += alpha*Y[i] where flag[i]==0. */
void orc_xplusay(double *X, size_t n, const int *flag, double alpha,
                 const double *Y);

/* Adjoint sensitivity of an element-wise Young's modulus scaling:
 * g[e] = -lam_e^T (dK_e/dE_e) u_e with K_e = E_e * K_e(E=1), i.e.
 * g[e] = -(1/E_e) lam_e^T K_e u_e evaluated as -lam_e^T K_e(mu1_e,lambda1_e) u_e
 * where (mu1_e, lambda1_e) are the Lame parameters at E=1.  Element matrix
 * from orc_emat_tet.  No reference counterpart (the reference has no adjoint). */
void orc_adjoint_grad_tet(double *grad, const double *mu1_e,
                          const double *lambda1_e, const double *xyz,
                          const unsigned *tet, size_t ntet, const double *u,
                          const double *lam);

/* ---- time-stepping drivers (fem_solidlinear.h:331-443) ------------------ */

/* fem_solidlinear.h:331-379 MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D,
 * element by element (ascending) through fem_solidlinear.cpp:437-490
 * EMat_SolidLinear_NewmarkBeta_MeshTet3D with is_initial_iter = true (what
 * the driver passes): emat = M_e + beta dt^2 K_e (SetEmat_LinearSolid3<4>,
 * fem_solidlinear.h:129-153; SetEmatConsistentMassTet, femutil.cpp:260-280,
 * with w0 = rho vol), eres = vol rho g / 4 - K_e u - M_e a - dt K_e v -
 * (0.5 dt^2) K_e a (AddEmatEvecScale3<4>, femutil.h:252-268, in that order);
 * merged with Merge<4,4,3,3>.  Per-element (myu_e, lambda_e); additive.
 * Returns 0, or 1 when a block is missing from the pattern. */
int orc_assemble_tet_newmark(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                             double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                             const double *lambda_e, double rho, const double *g, double dt,
                             double beta_newmark, const double *xyz, const unsigned *tet,
                             size_t ntet, const double *disp, const double *velo,
                             const double *acc);

/* fem_solidlinear.h:381-443 MergeLinSys_SolidLinear_BEuler_MeshTet3D:
 * K_e (SetEmat_LinearSolid3<4>), eres = vol rho g / 4 - K_e (u + dt v)
 * (three subtractions per node pair), then rho vol / 4 / dt^2 on the diagonal
 * entries of the diagonal blocks; vec_b += eres / dt.  Per-element
 * (myu_e, lambda_e); additive.  Returns 0, or 1 on a missing block. */
int orc_assemble_tet_beuler(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                            double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                            const double *lambda_e, double rho, const double *g, double dt,
                            const double *xyz, const unsigned *tet, size_t ntet,
                            const double *disp, const double *velo);

/* vecxitrsol.cpp:312-325 XPlusAYBZ: X += alpha Y + beta Z where flag == 0. */
void orc_xplusaybz(double *X, size_t n, const int *flag, double alpha, const double *Y,
                   double beta, const double *Z);
/* vecxitrsol.cpp:327-342 XPlusAYBZCW: X += alpha Y + beta Z + gamma W. */
void orc_xplusaybzcw(double *X, size_t n, const int *flag, double alpha, const double *Y,
                     double beta, const double *Z, double gamma, const double *W);

/* ---- stiffness warping (603_Fem3DStiffwarp) --------------------------------- */

/* vecxitrsol.h:350-437 Solve_PBiCGStab with the ILU(0) factors; history =
 * relative residual after each iteration (or |b| alone when |b|^2 < 1e-60).
 * Returns the history length. */
size_t orc_solve_pbicgstab(double *r, double *x, double tol, unsigned max_niter,
                           unsigned nblk, const unsigned *a_colind, const unsigned *a_rowptr,
                           const double *a_vcrs, const double *a_vdia,
                           const unsigned *l_colind, const unsigned *l_rowptr,
                           const unsigned *l_diaind, const double *l_vcrs,
                           const double *l_vdia, double *hist, size_t hist_cap);

/* svd3.cpp:114-305 GetRotPolarDecomp (EigenSym3 Jacobi sweeps on m^T m,
 * SortEigen3, Svd3's U, R = U V^T).  m, R row-major 3x3. */
void orc_rot_polar_decomp(double R[9], const double m[9], int nitr);

/* 603_Fem3DStiffwarp/main.cpp:80-106 RotationAtMeshPoints: R[9*np]. */
void orc_rotation_at_points(double *R, const double *xyz, const double *disp,
                            const unsigned *psup_ind, const unsigned *psup, size_t np);

/* fem_solidlinear.h:445-530 MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D
 * (per-element Lame pair; node rotations R[9*np]).  Returns 0, or 1 on a
 * missing block. */
int orc_assemble_tet_stiffwarp(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                               double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                               const double *lambda_e, double rho, const double *g, double dt,
                               const double *xyz, const unsigned *tet, size_t ntet,
                               const double *disp, const double *velo, const double *R);

#ifdef __cplusplus
}
#endif

#endif
