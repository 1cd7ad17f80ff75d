/*
 * dfm2_hip.h -- C ABI of the MI355X (gfx950) implementation of DelFEM2's
 * implicit tet-FEM solve path:
 *
 *   tet assembly -> 3x3 block-CSR CMatrixSparse -> CPreconditionerILU ILU(0)
 *   -> Solve_PCG
 *
 * Plain pointers and sizes only.  Unless stated otherwise, array arguments
 * named *_dev are device pointers and the others are host pointers; a
 * device argument that is pageable host memory (where the reference took a
 * host array) is rejected with DFM2H_ERR_INVALID before anything runs.
 * Every call is ordered on the calling thread's stream (dfm2h_set_stream);
 * calls that return host values synchronise that stream.  A handle is not
 * thread-safe; separate handles may be used from separate threads at the
 * same time, each thread on its own stream.  Every function returns a
 * status code (DFM2H_OK == 0); dfm2h_last_error() describes the last
 * failure of the calling thread.  No C++ exception crosses this boundary.
 *
 * Citations name the reference function each entry point replaces, relative
 * to the reference tree's include/delfem2/.
 *
 * Layout (same as CMatrixSparse<double> with 3x3 blocks):
 *   colind[nblk+1]  row-start offsets        (reference member col_ind_)
 *   rowptr[ncrs]    block column indices     (reference member row_ptr_)
 *   val_crs[9*ncrs] off-diagonal blocks, row-major 3x3
 *   val_dia[9*nblk] diagonal blocks, row-major 3x3
 *   vectors         AoS v[3*node + d]
 */
#ifndef DFM2_HIP_H
#define DFM2_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DFM2H_OK 0
#define DFM2H_ERR_INVALID 1       /* bad handle, size or argument            */
#define DFM2H_ERR_HIP 2           /* a HIP runtime call failed               */
#define DFM2H_ERR_PATTERN 3       /* Merge: block missing from the pattern   */
#define DFM2H_ERR_SINGULAR 4      /* Decompose: more than 10 singular pivots */
#define DFM2H_ERR_NOT_CONVERGED 5 /* Krylov solve hit max_nitr               */
#define DFM2H_ERR_NAN 6           /* residual became NaN / Inf               */
#define DFM2H_ERR_STATE 7         /* call out of order (e.g. no pattern yet) */

typedef struct dfm2h_bsr_s *dfm2h_bsr;         /* CMatrixSparse<double>, 3x3 */
typedef struct dfm2h_ilu_s *dfm2h_ilu;         /* CPreconditionerILU<double> */
typedef struct dfm2h_tetmesh_s *dfm2h_tetmesh; /* tet connectivity on device */
typedef struct dfm2h_batch_s *dfm2h_batch;     /* many meshes, one solver     */

/* ---- runtime ------------------------------------------------------------ */
/* Selects the HIP device for this thread.  No reference counterpart (the
 * reference is host-only). */
int dfm2h_init(int device);
/* The stream every later call of the calling thread is ordered on (a
 * hipStream_t; NULL = default stream); other threads keep their own.  No
 * reference counterpart. */
int dfm2h_set_stream(void *hip_stream);
void *dfm2h_get_stream(void);
int dfm2h_synchronize(void);
const char *dfm2h_last_error(void);
/* Device memory helpers for C/C++ callers without their own allocator. */
int dfm2h_malloc(void **ptr_dev, size_t bytes);
int dfm2h_free(void *ptr_dev);
int dfm2h_memcpy_htod(void *dst_dev, const void *src, size_t bytes);
int dfm2h_memcpy_dtoh(void *dst, const void *src_dev, size_t bytes);
int dfm2h_memcpy_dtod(void *dst_dev, const void *src_dev, size_t bytes);

/* ---- pattern (host) ------------------------------------------------------ */
/* msh_topology_uniform.cpp:301-316 JArray_PSuP_MeshElem followed by
 * jagarray.cpp:72-92 JArray_Sort, for 4-node tets.  psup_ind[np+1] is
 * filled; psup (nullable, capacity psup_cap) receives the sorted one-ring
 * without self; *npsup receives its length. */
int dfm2h_psup_tet(const unsigned *tet, size_t ntet, size_t np,
                   unsigned *psup_ind, unsigned *psup, size_t psup_cap,
                   size_t *npsup);
/* Node relabelling for the level schedule (no reference counterpart: the
 * reference factorises in the caller's node order, where a voxel cube of n^3
 * hexes has 3n+1 dependency levels per triangular sweep).  From a symmetric
 * one-ring pattern without self entries (as dfm2h_psup_tet builds it),
 * colours the nodes so that no two neighbours share a colour (first fit in
 * DSATUR order) and writes new2old[np]: the nodes colour by colour,
 * ascending old index inside a colour.  A mesh relabelled by it (node k of
 * the new mesh = node new2old[k] of the old) gives ILU(0) at most *ncolor
 * levels per sweep -- 4 on a Kuhn-split voxel grid.  The relabelled system is
 * the same equations; ILU(0)'s fill pattern, and so the iteration count,
 * depend on the order.  Host arrays; DFM2H_ERR_INVALID for a pattern that is
 * not symmetric, has self entries or entries out of range. */
int dfm2h_multicolor_order(size_t np, const unsigned *psup_ind, const unsigned *psup,
                           unsigned *new2old, unsigned *ncolor);

/* ---- CMatrixSparse<double> (ls_block_sparse.h) ---------------------------- */
/* ls_block_sparse.h:45-55 Initialize(nblk, len, is_dia); len must be 3 and
 * is_dia 1 (the tet path). */
int dfm2h_bsr_create(dfm2h_bsr *out, size_t nblk, unsigned len, int is_dia);
int dfm2h_bsr_destroy(dfm2h_bsr A);
/* ls_block_sparse.h:72-85 SetPattern(colind, ncolind, rowptr, nrowptr);
 * host arrays, copied.  May be called once per matrix. */
int dfm2h_bsr_set_pattern(dfm2h_bsr A, const unsigned *colind, size_t ncolind,
                          const unsigned *rowptr, size_t nrowptr);
/* ls_block_sparse.h:92-105 setZero. */
int dfm2h_bsr_set_zero(dfm2h_bsr A);
/* Host <-> device copies of val_crs_ / val_dia_ (the reference exposes them
 * as public host vectors; here they live in HBM).  Either pointer may be NULL. */
int dfm2h_bsr_upload(dfm2h_bsr A, const double *val_crs, const double *val_dia);
int dfm2h_bsr_download(dfm2h_bsr A, double *val_crs, double *val_dia);
/* Sizes and the device arrays (nullable outputs).  The arrays are for
 * reading; a caller that writes through them must call
 * dfm2h_bsr_values_changed afterwards, or later calls keep using copies
 * packed from the old values. */
int dfm2h_bsr_info(dfm2h_bsr A, size_t *nblk, size_t *ncrs,
                   double **val_crs_dev, double **val_dia_dev);
/* Declares that the values were written outside the library (through the
 * arrays of dfm2h_bsr_info).  No reference counterpart (the reference's
 * members are plain host vectors). */
int dfm2h_bsr_values_changed(dfm2h_bsr A);
/* ls_block_sparse.h:260-310 Merge<nrow, ncol, 3, 3, double>(A, ip_row,
 * ip_col, emat, merge_buffer): adds the element blocks emat[nrow][ncol][3][3]
 * (host, row-major; ip_row / ip_col host) into A, row by row: block (irow,
 * jcol) goes to stored slot (ip_row[irow], ip_col[jcol]), or -- when the two
 * nodes are the same -- emat[irow][irow] goes to the diagonal block, as
 * there.  The additions are staged on the host and applied on the device,
 * in call order per block, before A's values are next used: bit-identical
 * to the reference's sequential `+=`.  Returns DFM2H_ERR_PATTERN when a slot
 * is missing; the blocks before it (in the reference's order) are still
 * added and the rest of the element is not, as the reference returns false
 * there. */
int dfm2h_bsr_merge(dfm2h_bsr A, unsigned nrow, const unsigned *ip_row,
                    unsigned ncol, const unsigned *ip_col, const double *emat);
/* ls_block_sparse.cpp:215-273 MatVec: y = alpha*A*x + beta*y (y is scaled
 * by beta first, unconditionally, then off-diagonal blocks in stored order,
 * then the diagonal block: bit-identical to the reference). */
int dfm2h_bsr_matvec(dfm2h_bsr A, double *y_dev, double alpha,
                     const double *x_dev, double beta);
/* ls_block_sparse.h:147-151 SetFixedBC(bc_flag) (-> cpp:397-472):
 * bc_flag_dev[3*nblk], nonzero = fixed DOF. */
int dfm2h_bsr_set_fixed_bc(dfm2h_bsr A, const int *bc_flag_dev);
/* vecxitrsol.cpp:286-296 setRHS_Zero: b[i]=0 where flag[i]!=iflag_nonzero. */
int dfm2h_set_rhs_zero(double *b_dev, const int *flag_dev, size_t n,
                       int iflag_nonzero);
/* vecxitrsol.cpp:226-237 XPlusAY: X[i] += alpha*Y[i] where flag[i]==0. */
int dfm2h_xplusay(double *X_dev, size_t n, const int *flag_dev, double alpha,
                  const double *Y_dev);

/* ---- tet assembly (fem_solidlinear.h) ------------------------------------ */
/* Element connectivity for assembly: tet[4*ntet] (host, copied) over np
 * nodes.  Builds the element-around-node lists in ascending element order
 * (msh_topology_uniform.cpp:19-51) that the gather assembly walks. */
int dfm2h_tetmesh_create(dfm2h_tetmesh *out, const unsigned *tet, size_t ntet,
                         size_t np);
/* The same pattern built on the device from the mesh handle (one thread per
 * node over its surrounding elements; msh_topology_uniform.cpp:301-316 +
 * jagarray.cpp:72-92, identical output).  Host outputs as dfm2h_psup_tet:
 * psup_ind[np+1] always, psup when psup_cap >= *npsup.  Any number of
 * elements around a node (up to 32 the candidates stay in LDS, beyond that
 * in a device buffer). */
int dfm2h_psup_tet_device(dfm2h_tetmesh M, unsigned *psup_ind, unsigned *psup,
                          size_t psup_cap, size_t *npsup);
int dfm2h_tetmesh_destroy(dfm2h_tetmesh M);
/* fem_solidlinear.h:192-237 MergeLinSys_SolidLinear_Static_MeshTet3D with
 * EMat_SolidLinear_Static_Tet (fem_solidlinear.cpp:409-435) merged by
 * Merge<4,4,3,3> (ls_block_sparse.h:260-310).  Additive into A and
 * vec_b_dev, bit-identical to the reference's ascending-element loop.
 * myu_e_dev / lambda_e_dev: per-element Lame parameters [ntet], or NULL to
 * use the scalars myu / lambda (the reference signature).  xyz_dev [3*np],
 * disp_dev [3*np] or NULL (zero displacement), g[3] host.
 * Returns DFM2H_ERR_PATTERN if an element edge is missing from A's pattern
 * (the reference's Merge returns false). */
int dfm2h_merge_linsys_solidlinear_static_tet(
    dfm2h_bsr A, double *vec_b_dev, const double *myu_e_dev,
    const double *lambda_e_dev, double myu, double lambda, double rho,
    const double *g, const double *xyz_dev, dfm2h_tetmesh M,
    const double *disp_dev);
/* fem_solidlinear.h:331-379 MergeLinSys_SolidLinear_NewmarkBeta_MeshTet3D:
 * per element (ascending) EMat_SolidLinear_NewmarkBeta_MeshTet3D
 * (fem_solidlinear.cpp:437-490) with is_initial_iter = true, as the driver
 * passes: M_e + beta dt^2 K_e into A (consistent mass, femutil.cpp:260-280)
 * and vol rho g / 4 - K_e u - M_e a - dt K_e v - (0.5 dt^2) K_e a into
 * vec_b; bit-identical to the reference.  gamma_newmark is unused by the
 * element (as in the reference).  disp / velo / acc_dev [3*np] (NULL: zero);
 * other arguments as for the static assembly. */
int dfm2h_merge_linsys_solidlinear_newmark_tet(
    dfm2h_bsr A, double *vec_b_dev, const double *myu_e_dev,
    const double *lambda_e_dev, double myu, double lambda, double rho,
    const double *g, double dt, double gamma_newmark, double beta_newmark,
    const double *xyz_dev, dfm2h_tetmesh M, const double *disp_dev,
    const double *velo_dev, const double *acc_dev);
/* fem_solidlinear.h:381-443 MergeLinSys_SolidLinear_BEuler_MeshTet3D:
 * K_e plus rho vol / 4 / dt^2 on the diagonal entries of the diagonal
 * blocks into A, (vol rho g / 4 - K_e (u + dt v)) / dt into vec_b;
 * bit-identical to the reference. */
int dfm2h_merge_linsys_solidlinear_beuler_tet(
    dfm2h_bsr A, double *vec_b_dev, const double *myu_e_dev,
    const double *lambda_e_dev, double myu, double lambda, double rho,
    const double *g, double dt, const double *xyz_dev, dfm2h_tetmesh M,
    const double *disp_dev, const double *velo_dev);
/* fem_solidlinear.h:445-530 MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D:
 * blocks R_i K_e,ij R_i^T (R_i the rotation of the row node, rot_dev[9*np]
 * row-major, e.g. from dfm2h_rotation_at_points) plus rho vol / 4 / dt^2 on
 * the diagonal entries of the diagonal blocks; vec_b += (vol rho g / 4 -
 * sum_j E_ij (P_j + u_j + dt v_j - R_i P_j)) / dt.  Bit-identical to the
 * reference for the same rotations.  The matrix is not symmetric: solve with
 * dfm2h_solve_pbicgstab. */
int dfm2h_merge_linsys_solidstiffwarp_beuler_tet(
    dfm2h_bsr A, double *vec_b_dev, const double *myu_e_dev,
    const double *lambda_e_dev, double myu, double lambda, double rho,
    const double *g, double dt, const double *xyz_dev, dfm2h_tetmesh M,
    const double *disp_dev, const double *velo_dev, const double *rot_dev);
/* 603_Fem3DStiffwarp/main.cpp:80-106 RotationAtMeshPoints: per node, the
 * polar-decomposition rotation (svd3.cpp:291-305 GetRotPolarDecomp, 100
 * Jacobi sweeps) of sum_j (p_j - p_i)(P_j - P_i)^T over the node's one-ring
 * (A's pattern), p = P + u.  rot_dev[9*nblk] row-major; disp_dev NULL =
 * zero.  Equal to the reference to rounding (device trigonometry). */
int dfm2h_rotation_at_points(dfm2h_bsr A, const double *xyz_dev, const double *disp_dev,
                             double *rot_dev);
/* vecxitrsol.cpp:312-325 XPlusAYBZ: X += alpha*Y + beta*Z where flag == 0. */
int dfm2h_xplusaybz(double *X_dev, size_t n, const int *flag_dev, double alpha,
                    const double *Y_dev, double beta, const double *Z_dev);
/* vecxitrsol.cpp:327-342 XPlusAYBZCW: X += alpha*Y + beta*Z + gamma*W where
 * flag == 0. */
int dfm2h_xplusaybzcw(double *X_dev, size_t n, const int *flag_dev, double alpha,
                      const double *Y_dev, double beta, const double *Z_dev,
                      double gamma, const double *W_dev);

/* ---- CPreconditionerILU<double> (ls_ilu_block_sparse.h) ------------------- */
int dfm2h_ilu_create(dfm2h_ilu *out);
int dfm2h_ilu_destroy(dfm2h_ilu P);
/* ls_ilu_block_sparse.cpp:1057-1090 SetPattern0(m) (also what
 * Initialize_ILUk(m, 0) does, :854-857): copies A's pattern, sorts each row,
 * finds the first upper slot of each row, and schedules the rows in
 * dependency levels for the parallel factorisation and sweeps. */
int dfm2h_ilu_set_pattern0(dfm2h_ilu P, dfm2h_bsr A);
/* ls_ilu_block_sparse.cpp:849-999 Initialize_ILUk: the ILU(k) pattern of A
 * (lev_fill 0: SetPattern0; -1: every fill).  The numeric CopyValue /
 * Decompose / sweeps and the fused solvers work on it unchanged. */
int dfm2h_ilu_set_pattern_k(dfm2h_ilu P, dfm2h_bsr A, int lev_fill);
/* block rows, stored off-diagonal blocks, and (nullable, nblk+1 entries)
 * the row starts of the ILU pattern (colInd; A's own for ILU(0)). */
int dfm2h_ilu_info(dfm2h_ilu P, size_t *nblk, size_t *ncrs, unsigned *colind);
/* ls_ilu_block_sparse.cpp:1009-1047 CopyValue(m). */
int dfm2h_ilu_copy_value(dfm2h_ilu P, dfm2h_bsr A);
/* ls_ilu_block_sparse.cpp:152-420 Decompose() (3x3 branch 286-356).  Rows
 * are factorised level by level, each row in the reference's IKJ order, so
 * the factors are bit-identical.  *ok = 1 (true), or 0 when more than 10
 * pivots were singular (|det| <= 1e-30): the reference then returns at its
 * 11th singular pivot in row order, and the factors are left as it leaves
 * them -- rows before that row factorised, that row eliminated but its U
 * part not scaled, later rows as CopyValue wrote them (restored from the
 * matrix CopyValue read; if that matrix was destroyed or changed since,
 * the later rows stay factorised) -- and *n_singular = 11.  Either output
 * may be NULL.  Returns DFM2H_OK in both cases. */
int dfm2h_ilu_decompose(dfm2h_ilu P, int *ok, int *n_singular);
/* ls_ilu_block_sparse.h:38-41 SolvePrecond(vec) = ForwardSubstitution
 * (cpp:544-573) then BackwardSubstitution (cpp:721-749), in place. */
int dfm2h_ilu_solve_precond(dfm2h_ilu P, double *vec_dev);
int dfm2h_ilu_forward(dfm2h_ilu P, double *vec_dev);
int dfm2h_ilu_backward(dfm2h_ilu P, double *vec_dev);
/* Factor values (valCrs, valDia) and the sorted pattern (rowPtr, m_diaInd);
 * host outputs, each nullable. */
int dfm2h_ilu_download(dfm2h_ilu P, unsigned *rowptr, unsigned *diaind,
                       double *val_crs, double *val_dia);
/* Number of forward / backward dependency levels of the schedule. */
int dfm2h_ilu_levels(dfm2h_ilu P, unsigned *n_forward, unsigned *n_backward);

/* ---- view_vectorx.h operator set (device vectors of n doubles) ------------ */
/* view_vectorx.h:104-109 Dot -> host double (deterministic tree order). */
int dfm2h_vec_dot(const double *x_dev, const double *y_dev, size_t n,
                  double *out);
/* No counterpart in the reference: dfm2h_vec_dot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and summation order, so
 * the result is bit-identical to dfm2h_vec_dot; it calls no hipMalloc /
 * hipFree.  Its scratch (2048 workgroup partials and the result, 16 KiB) is
 * a static device array that each GPU holds until the process ends; every
 * call on one GPU shares it, so calls that overlap on one GPU (other
 * threads or streams) race.  Any n is accepted (the partials never exceed
 * the scratch). */
int dfm2h_vec_dot_noalloc(const double *x_dev, const double *y_dev, size_t n,
                          double *out);
/* view_vectorx.h:61-71 AddScaledVec: y += alpha*x. */
int dfm2h_vec_add_scaled(double *y_dev, double alpha, const double *x_dev,
                         size_t n);
/* view_vectorx.h:73-83 ScaleAndAddVec: y = beta*y + x. */
int dfm2h_vec_scale_and_add(double *y_dev, double beta, const double *x_dev,
                            size_t n);
/* view_vectorx.h:44-47 deep-copy operator= and :49-51 setZero. */
int dfm2h_vec_copy(double *dst_dev, const double *src_dev, size_t n);
int dfm2h_vec_set_zero(double *x_dev, size_t n);

/* ---- Solve_PBiCGStab (vecxitrsol.h:350-437) ------------------------------ */
/* ILU-preconditioned BiCGStab for non-symmetric systems (the stiffness-warped
 * backward-Euler matrix): x_dev set to zero, r_dev holds b on entry and the
 * final residual on return.  hist receives the relative residual after each
 * iteration (the reference's history: no entry for the start; a single
 * entry |b| when |b|^2 < 1e-60); at most max_niter - 1 iterations, as there.
 * The reference's recurrences operation for operation on the device; the
 * dot products (deterministic tree order) are read by the host. */
int dfm2h_solve_pbicgstab(dfm2h_bsr A, dfm2h_ilu P, double *r_dev, double *x_dev,
                          double tol, unsigned max_niter, double *hist, size_t hist_cap,
                          size_t *n_hist);

/* ---- fused Solve_PCG (lsitrsol.h:80-148) ---------------------------------- */
/* Solve_PCG(r, x, Pr, p, tol, max_nitr, A, ilu) with the preconditioner P
 * (already decomposed from A).  r_dev holds b on entry and the final
 * residual on return; x_dev receives the solution (set to zero first).
 * hist (host, capacity hist_cap, nullable) receives the residual-norm
 * history, absolute norms, entry 0 = |b|; *n_hist its length (as the
 * reference's returned vector).  The whole iteration runs on the device
 * (captured HIP graph, scalars kept in HBM); the host only polls the
 * convergence flag between chunks of iterations.  Returns DFM2H_OK also
 * when max_nitr is reached (as the reference); *n_hist tells. */
int dfm2h_solve_pcg(dfm2h_bsr A, dfm2h_ilu P, double *r_dev, double *x_dev,
                    double tol, unsigned max_nitr, double *hist,
                    size_t hist_cap, size_t *n_hist);

/* Mixed precision (no reference counterpart; the reference has no float
 * ILU or PCG): an FP32 ILU(0)-PCG (FP32 BSR values, FP32 factors rounded
 * from P's FP64 factors) inside FP64 iterative refinement by reliable
 * updates.  Whenever the FP32 residual has fallen by the factor `delta`
 * (0 < delta < 1) since the last refinement step, or looks converged, the
 * FP32 correction is folded into the FP64 solution x, the residual
 * r = b - A x is recomputed in FP64 (FP64 values of A), the stop test
 * |r|/|b| < tol is made in FP64, and the FP32 iteration continues from the
 * true residual with its search direction kept (no Krylov restart).
 * The FP32 copy of A is held in difference form: the off-diagonal blocks
 * multiply x_j - x_i, and the diagonal slot holds the row's block sum
 * D_i + sum_j A_ij (summed in FP64, rounded once), so that the rounding of
 * the stored blocks costs no extra iterations on smooth vectors.
 * When a reliable update finds that 500 inner iterations have passed since
 * the last one that halved the smallest FP64 residual so far (the FP32
 * matrix too coarse for the system, e.g. element stiffnesses spread over ~7
 * decades), the FP32 iteration stops and the
 * FP64 solver (dfm2h_solve_pcg, up to max_inner iterations) finishes on
 * the correction: x += A^-1 (b - A x); dfm2h_pcg_mixed_fell_back tells.
 * b_dev is not modified; x_dev receives the solution.  hist receives the
 * FP64 residual norms (entry 0 = |b|, then one per refinement step, then,
 * after a fallback, the FP64 solver's history, whose norms are those of
 * b - A x); n_inner the number of FP32 iterations plus the FP64 solver's.
 * max_outer caps the refinement steps, max_inner the FP32 iterations. */
int dfm2h_solve_pcg_mixed(dfm2h_bsr A, dfm2h_ilu P, const double *b_dev,
                          double *x_dev, double tol, double delta,
                          unsigned max_outer, unsigned max_inner,
                          double *hist, size_t hist_cap, size_t *n_hist,
                          size_t *n_inner);

/* Whether P's last dfm2h_solve_pcg_mixed stalled and was finished by the
 * FP64 solver (no reference counterpart): *fell_back = 1 or 0. */
int dfm2h_pcg_mixed_fell_back(dfm2h_ilu P, int *fell_back);

/* Options of the mixed solver (no reference counterpart).
 * fp64_vectors = 0 (default): r, x, p, Pr in FP32 (the solution kept as
 * x64 + x32), dot products accumulated in FP64; = 1: the vectors in FP64,
 * the matrix and factor values still streamed in FP32 (widened before the
 * products).  update_period > 0 forces a reliable update after that many
 * inner iterations without one (0: only the delta rule).  Defaults: FP32
 * vectors, period 100 (measured on MI355X with delta = 0.1: the inner
 * iteration count of the 64^3 cube stays at the FP64 solver's with 9
 * reliable updates instead of 13 at period 50 -- 58.0 against 58.5 ms --,
 * and the 128^3 cube to 1e-10 takes 980 against 987 ms; period 200 and no
 * forced updates need more iterations there). */
int dfm2h_pcg_set_mixed_options(dfm2h_ilu P, int fp64_vectors, unsigned update_period);
/* HIP-graph chunking of the fused solvers (no reference counterpart): the
 * iterations are captured in graphs of `large` and of `small` iterations;
 * small ones run first and once the residual's recent rate of decrease
 * predicts the end within two large ones (fewer iterations run, as
 * immediate returns, after convergence), large ones in between.  Defaults
 * 16 and 2; small == large gives a fixed chunk.  The chunking changes no
 * result.  Requires 1 <= small <= large <= 1024. */
int dfm2h_pcg_set_graph_chunks(dfm2h_ilu P, unsigned large, unsigned small);
/* Where the fused solvers' beta step runs (no reference counterpart): the
 * step (stop test, beta = (r, Pr)_new / (r, Pr)_old, and the mixed solver's
 * reliable-update decision) as a one-workgroup kernel over the partial sums
 * of the r/x update and the sweeps (mode 0), or inside the p update, every
 * workgroup of which then sums all those partials itself (mode 1: one
 * launch fewer per iteration, partial reads that grow with the square of
 * the system size).  -1 (the default): folded while those reads stay below
 * 24 MiB per iteration (823,875 DOF: 13 MiB).  Both forms give the same
 * bits. */
int dfm2h_pcg_set_scalar_fold(dfm2h_ilu P, int mode);
/* Launch shape of the fused solvers (no reference counterpart): shapes
 * 0..5 set the triangular sweeps' {wavefronts per 21-row slice, blocks per
 * load round} from { {1,8}, {1,4}, {2,8}, {2,4}, {1,16}, {2,16} }; with two
 * wavefronts per slice each takes every other block of a row and the two
 * sums are added (a different rounding order from the reference's loops;
 * shapes 0, 1, 4 keep it).  Shapes 6..11 are shapes 0..5 with the slice
 * kernels' workgroups numbered so that each XCD (own L2) runs one contiguous
 * run of rows; that changes no sum, so shape s + 6 gives the bits of shape
 * s.  -1 (the default): the first solve of each precision times one SpMV
 * plus one sweep pair for every shape (FP64: shapes 0, 1, 4, 6, 7, 10 only)
 * and keeps the fastest, per (floor(log2 ndof), solver kind) for the
 * process (and in the file named by DFM2H_TUNE_CACHE, if set).  fp32 = 0
 * names the FP64 solver, else the mixed solver with its current vector
 * precision (each has its own shape).  get_shape: current shape, its name,
 * and the autotuner's ms per round of every shape (0 where not measured);
 * n_shapes = 12. */
int dfm2h_pcg_set_shape(dfm2h_ilu P, int fp32, int shape);
int dfm2h_pcg_get_shape(dfm2h_ilu P, int fp32, int *shape, const char **name,
                        double *tune_ms, size_t cap, size_t *n_shapes);
/* Measurement (no reference counterpart): average time in ms of each kernel
 * class inside the fused iteration, from HIP events recorded between
 * consecutive kernels on the solver's own stream over nrep iterations run
 * right after a solve of the same precision (fp32 = 1: the mixed solver's
 * FP32 iteration); convergence and refinement are disabled meanwhile.
 * ms[8] = { SpMV, r/x update with forward level 0, forward levels >= 1,
 * backward levels >= 1 (backward level 0 has no work: its (r, Pr) share is
 * summed in the forward sweep), stop test + beta step + p update, alpha
 * step, the SpMV again timed
 * over nrep launches back to back (no events between kernels: its duration
 * plus the dispatch gap), 1 if the beta step ran folded else 0 } (the
 * FP32 reliable-update residual runs inside the SpMV when due, its decision
 * inside the beta step).  The kernels run on the solver's scratch vectors;
 * nothing the caller owns is touched. */
int dfm2h_pcg_kernel_times(dfm2h_bsr A, dfm2h_ilu P, int fp32, unsigned nrep,
                           double *ms);
/* Measurement (no reference counterpart): read + write bytes per second
 * (GB/s) of a plain device copy of `bytes` bytes (16-byte vectors, one per
 * thread), averaged over `reps` launches on the library stream -- the
 * achievable counterpart of the HBM peak for the kernels' rooflines. */
int dfm2h_measure_copy(size_t bytes, unsigned reps, double *gbs);

/* ---- adjoint (no reference counterpart) --------------------------------- */
/* Material sensitivity of an objective J(u) with K(E) u = f and
 * K(E) = sum_e E_e K_e^1, K_e^1 the element stiffness at the Lame pair
 * (myu1_e, lambda1_e) of unit modulus (fem_solidlinear.cpp:409-435):
 *   grad_dev[e] = -lam_e^T K_e^1 u_e      for every element e,
 * where lam_dev solves K lam = dJ/du (zero on fixed DOFs; for the compliance
 * J = f.u, lam = u).  Device arrays: xyz [3*np], myu1_e / lambda1_e / grad
 * [ntet], u / lam [3*np]. */
int dfm2h_adjoint_grad_tet(dfm2h_tetmesh M, const double *xyz_dev,
                           const double *myu1_e_dev, const double *lambda1_e_dev,
                           const double *u_dev, const double *lam_dev,
                           double *grad_dev);

/* ---- batched solve (no reference counterpart) --------------------------- */
/* Many independent meshes solved together.  A and P hold the union system:
 * mesh m owns nodes [node_offset[m], node_offset[m+1]) and no block of A
 * couples two meshes (checked: DFM2H_ERR_PATTERN).  A and P are assembled,
 * given fixed BCs, copied and decomposed with the ordinary calls; the batch
 * keeps its own mesh-aligned packed plans, built here from the patterns
 * (rebuild the batch after setting a new pattern: DFM2H_ERR_STATE).  Values
 * are re-packed whenever A or P changed.  node_offset: host, nmesh+1
 * entries. */
int dfm2h_batch_create(dfm2h_batch *out, dfm2h_bsr A, dfm2h_ilu P,
                       const unsigned *node_offset, size_t nmesh);
int dfm2h_batch_destroy(dfm2h_batch B);
/* nslices: wavefront slices of the matrix plan; n_forward / n_backward:
 * dependency levels of the union factors (= sweep launches per iteration). */
int dfm2h_batch_info(dfm2h_batch B, size_t *nmesh, size_t *nslices,
                     unsigned *n_forward, unsigned *n_backward);
/* One ILU(0)-PCG per mesh, all meshes in the same kernels (Solve_PCG of
 * lsitrsol.h:80-148 per mesh).  b_dev, x_dev: union vectors [3*np], device.
 * fp32 = 0: FP64 throughout, stop when |r_k| / |r_0| < tol per mesh (the
 * reference's test); fp32 = 1: FP32 SpMV / sweeps / vectors inside FP64
 * reliable updates (x64 += x32, r = b - A x64 in FP64 whenever the FP32
 * residual fell by delta), stop when the FP64 |b - A x| / |b| < tol.
 * max_nitr caps the (inner) iterations of each mesh.  The FP32 matrix copy
 * is held in the difference form of dfm2h_solve_pcg_mixed.  A mesh whose
 * FP32 iteration stalls (500 inner iterations without a reliable update that
 * halved its smallest FP64 residual) stops, and after the batch the FP64 form
 * finishes every such mesh on its correction, x += A^-1 (b - A x), stopping
 * when |b - A x| / |b| < tol (dfm2h_batch_fell_back tells which).  Optional
 * host outputs per mesh: iters (inner iterations; after an FP64 finish, the
 * FP32 ones plus the FP64 ones), relres (last relative residual tested),
 * state (1 converged, 0 iteration cap, -1 NaN).  Returns DFM2H_ERR_NAN if any
 * mesh produced NaN / Inf; meshes that hit the cap are reported in state.
 * The first solve of a (DOF bucket, precision) with no variant set runs the
 * autotuner (dfm2h_batch_tune). */
/* Per mesh (host, nmesh entries, no reference counterpart): 1 when the last
 * dfm2h_batch_solve in the FP32 form finished that mesh with the FP64 form
 * after its FP32 iteration stalled, else 0. */
int dfm2h_batch_fell_back(dfm2h_batch B, int *flags);
int dfm2h_batch_solve(dfm2h_batch B, const double *b_dev, double *x_dev,
                      double tol, int fp32, double delta, unsigned max_nitr,
                      unsigned *iters, double *relres, int *state);
/* Autotuner: times a 32-iteration solve of every launch-shape variant
 * (wavefronts per workgroup, slices per wavefront, blocks per load round of
 * the SpMV and of the sweeps) on this batch, keeps the fastest for this
 * batch and for its (floor(log2 ndof), precision) key; ms[v] = ms per
 * iteration of variant v (ms_cap entries at most).  The key -> variant map
 * lives for the process; when the environment variable DFM2H_TUNE_CACHE
 * names a file it is read from and appended to that file as well. */
int dfm2h_batch_tune(dfm2h_batch B, int fp32, int *best, double *ms,
                     size_t ms_cap, size_t *n_variants);
/* FP32 form: force a reliable update of a mesh after update_period inner
 * iterations without one (0: only the delta rule; default 25). */
int dfm2h_batch_set_update_period(dfm2h_batch B, unsigned update_period);
/* variant -1 = autotune on the next solve; 0 .. n_variants-1 = fixed. */
int dfm2h_batch_set_variant(dfm2h_batch B, int fp32, int variant);
int dfm2h_batch_get_variant(dfm2h_batch B, int fp32, int *variant,
                            const char **name);

#ifdef __cplusplus
}
#endif

#endif /* DFM2_HIP_H */
