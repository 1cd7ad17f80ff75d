// bsr.hip -- CMatrixSparse<double> (3x3 blocks, separate diagonal) on the
// device: pattern, values, MatVec, Dirichlet rows/columns, and the gather
// form of the tet assembly.
#include "handles.hpp"
#include "tet.hpp"

namespace dfm2h {

struct AsmArgs {
  unsigned nblk;
  size_t ncrs;
  const unsigned *colind, *rowptr, *rowof;
  double *vcrs, *vdia, *vec_b;
  const double *myu_e, *lambda_e;
  double myu, lambda, rho, g0, g1, g2;
  const double *xyz, *disp;
  const uint4 *tet;
  const unsigned *elsup_ind, *elsup;
  int *missing;  // set when an element edge has no slot in the pattern
};

// Gather assembly, one thread per destination block: off-diagonal slot s
// (t < ncrs) or the diagonal block and right-hand side of row t-ncrs.  Each
// thread walks the elements around its row node in ascending element order
// and adds their contribution to its block; the additions therefore happen
// in exactly the order of the reference's sequential element loop with
// Merge<4,4,3,3>, and no two threads write the same address (no atomics).
__global__ void __launch_bounds__(256) k_assemble_tet_gather(AsmArgs A) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool diag = t >= A.ncrs;
  if (t >= A.ncrs + A.nblk) return;
  const unsigned i = diag ? (unsigned)(t - A.ncrs) : A.rowof[t];
  const unsigned j = diag ? i : A.rowptr[t];
  double *dst = diag ? A.vdia + (size_t)i * 9 : A.vcrs + t * 9;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = dst[q];
  double bacc[3] = {0, 0, 0};
  if (diag) {
    bacc[0] = A.vec_b[(size_t)i * 3 + 0];
    bacc[1] = A.vec_b[(size_t)i * 3 + 1];
    bacc[2] = A.vec_b[(size_t)i * 3 + 2];
  }
  const unsigned q1 = A.elsup_ind[i + 1];
  for (unsigned q = A.elsup_ind[i]; q < q1; ++q) {
    const unsigned e = A.elsup[q];
    const uint4 n = A.tet[e];
    const int li = local_index(n, i);
    const int lj = diag ? li : local_index(n, j);
    if (lj < 0) continue;
    double P[4][3];
    load_tet(P, A.xyz, n);
    TetGeo G;
    tet_geometry(G, P);
    const double myu = A.myu_e ? A.myu_e[e] : A.myu;
    const double lam = A.lambda_e ? A.lambda_e[e] : A.lambda;
    const double w = G.vol;
    double di[3], dj[3];
    pick_row(di, G, li);
    pick_row(dj, G, lj);
    double blk[9];
    emat_block(blk, di, dj, w, myu, lam);
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] += blk[k];
    if (diag) {
      // eres = vol*rho*g*0.25 - K_e u_e (fem_solidlinear.h:214-221 and
      // AddEmatEvecScale3 femutil.h:252-268, scale -1), then vec_b += eres
      const double vol = G.vol;
      double er0 = vol * A.rho * A.g0 * 0.25;
      double er1 = vol * A.rho * A.g1 * 0.25;
      double er2 = vol * A.rho * A.g2 * 0.25;
      const unsigned id[4] = {n.x, n.y, n.z, n.w};
#pragma unroll
      for (int jn = 0; jn < 4; ++jn) {
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
        if (A.disp) {
          u0 = A.disp[(size_t)id[jn] * 3 + 0];
          u1 = A.disp[(size_t)id[jn] * 3 + 1];
          u2 = A.disp[(size_t)id[jn] * 3 + 2];
        }
        double kb[9];
        emat_block(kb, di, G.d[jn], w, myu, lam);
        er0 += -1. * (kb[0] * u0 + kb[1] * u1 + kb[2] * u2);
        er1 += -1. * (kb[3] * u0 + kb[4] * u1 + kb[5] * u2);
        er2 += -1. * (kb[6] * u0 + kb[7] * u1 + kb[8] * u2);
      }
      bacc[0] += er0;
      bacc[1] += er1;
      bacc[2] += er2;
      // every neighbour of i in this element must own a slot in row i
      const unsigned r0 = A.colind[i], r1 = A.colind[i + 1];
#pragma unroll
      for (int jn = 0; jn < 4; ++jn) {
        const unsigned col = id[jn];
        if (col == i) continue;
        unsigned lo = r0, hi = r1;
        while (lo < hi) {
          const unsigned mid = (lo + hi) >> 1;
          if (A.rowptr[mid] < col) lo = mid + 1; else hi = mid;
        }
        bool found = lo < r1 && A.rowptr[lo] == col;
        if (!found)  // pattern rows need not be sorted: fall back to a scan
          for (unsigned s = r0; s < r1 && !found; ++s) found = A.rowptr[s] == col;
        if (!found) atomicOr(A.missing, 1);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) dst[q] = acc[q];
  if (diag) {
    A.vec_b[(size_t)i * 3 + 0] = bacc[0];
    A.vec_b[(size_t)i * 3 + 1] = bacc[1];
    A.vec_b[(size_t)i * 3 + 2] = bacc[2];
  }
}

// SetFixedBC (ls_block_sparse.cpp:397-472): entries in a fixed row or a
// fixed column become 0; the diagonal entry of a fixed DOF becomes 1.
__global__ void k_set_fixed_bc(unsigned nblk, size_t ncrs, const unsigned *__restrict__ rowof,
                               const unsigned *__restrict__ rowptr, double *vcrs, double *vdia,
                               const int *__restrict__ flag) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncrs + nblk) return;
  const bool diag = t >= ncrs;
  const unsigned i = diag ? (unsigned)(t - ncrs) : rowof[t];
  const unsigned j = diag ? i : rowptr[t];
  double *v = diag ? vdia + (size_t)i * 9 : vcrs + t * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const bool fr = flag[(size_t)i * 3 + a] != 0;
      const bool fc = flag[(size_t)j * 3 + b] != 0;
      if (diag) {
        if (fr || fc) v[a * 3 + b] = (a == b && fr) ? 1.0 : 0.0;
      } else {
        if (fr || fc) v[a * 3 + b] = 0.0;
      }
    }
}

__global__ void k_set_rhs_zero(double *b, const int *flag, size_t n, int nz) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i] != nz) b[i] = 0;
}

__global__ void k_xplusay(double *X, size_t n, const int *flag, double alpha, const double *Y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i] == 0) X[i] += alpha * Y[i];
}

__global__ void k_fill_rowof(unsigned nblk, const unsigned *colind, unsigned *rowof) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblk) return;
  for (unsigned s = colind[i]; s < colind[i + 1]; ++s) rowof[s] = i;
}

}  // namespace dfm2h

using namespace dfm2h;

static void need_pattern(dfm2h_bsr A) {
  DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
  DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_STATE, "SetPattern has not been called");
}

extern "C" {

int dfm2h_bsr_create(dfm2h_bsr *out, size_t nblk, unsigned len, int is_dia) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    DFM2H_REQUIRE(len == 3, DFM2H_ERR_INVALID, "only 3x3 blocks are supported (len must be 3)");
    DFM2H_REQUIRE(is_dia, DFM2H_ERR_INVALID, "the separate diagonal (is_dia) is required");
    DFM2H_REQUIRE(nblk < (1ull << 31), DFM2H_ERR_INVALID, "too many block rows");
    auto *A = new dfm2h_bsr_s;
    A->nblk = (unsigned)nblk;
    try {
      A->vdia.alloc(nblk * 9);
      A->vdia.zero();
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}

int dfm2h_bsr_destroy(dfm2h_bsr A) {
  return guard(__func__, [&] { delete A; });
}

int dfm2h_bsr_set_pattern(dfm2h_bsr A, const unsigned *colind, size_t ncolind,
                          const unsigned *rowptr, size_t nrowptr) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    DFM2H_REQUIRE(!A->has_pattern, DFM2H_ERR_STATE, "SetPattern may be called once");
    DFM2H_REQUIRE(colind && ncolind == (size_t)A->nblk + 1, DFM2H_ERR_INVALID,
                  "ncolind must be nblk+1");
    DFM2H_REQUIRE(colind[0] == 0 && colind[A->nblk] == nrowptr, DFM2H_ERR_INVALID,
                  "colind must start at 0 and end at nrowptr");
    DFM2H_REQUIRE(rowptr || nrowptr == 0, DFM2H_ERR_INVALID, "null rowptr");
    for (unsigned i = 0; i < A->nblk; ++i)
      DFM2H_REQUIRE(colind[i] <= colind[i + 1], DFM2H_ERR_INVALID, "colind not monotone");
    for (size_t s = 0; s < nrowptr; ++s)
      DFM2H_REQUIRE(rowptr[s] < A->nblk, DFM2H_ERR_INVALID, "column index out of range");
    Pattern &P = A->pat;
    P.nblk = A->nblk;
    P.ncrs = nrowptr;
    P.colind.assign(colind, colind + ncolind);
    P.rowptr.assign(rowptr, rowptr + nrowptr);
    P.colind_d.upload(P.colind);
    P.rowptr_d.upload(P.rowptr);
    P.rowof_d.alloc(nrowptr);
    if (A->nblk && nrowptr) {
      hipLaunchKernelGGL(k_fill_rowof, dim3(ceil_div(A->nblk, 256)), dim3(256), 0, stream(),
                         A->nblk, P.colind_d.p, P.rowof_d.p);
      check_launch("k_fill_rowof");
    }
    A->vcrs.alloc(nrowptr * 9);
    A->vcrs.zero();
    {
      std::vector<unsigned> all(A->nblk), s0(A->nblk), s1(A->nblk);
      for (unsigned i = 0; i < A->nblk; ++i) {
        all[i] = i;
        s0[i] = colind[i];
        s1[i] = colind[i + 1];
      }
      A->plan.build({all}, s0, s1, P.rowptr, true);
    }
    A->has_pattern = true;
    A->pattern_serial++;
    A->version++;
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_bsr_set_zero(dfm2h_bsr A) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    A->vcrs.zero();
    A->vdia.zero();
    A->version++;
  });
}

int dfm2h_bsr_upload(dfm2h_bsr A, const double *vcrs, const double *vdia) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    if (vcrs) {
      need_pattern(A);
      A->vcrs.upload(vcrs, A->pat.ncrs * 9);
    }
    if (vdia) A->vdia.upload(vdia, (size_t)A->nblk * 9);
    A->version++;
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_bsr_download(dfm2h_bsr A, double *vcrs, double *vdia) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    if (vcrs) {
      need_pattern(A);
      A->vcrs.download(vcrs, A->pat.ncrs * 9);
    }
    if (vdia) A->vdia.download(vdia, (size_t)A->nblk * 9);
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_bsr_info(dfm2h_bsr A, size_t *nblk, size_t *ncrs, double **vcrs, double **vdia) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    if (nblk) *nblk = A->nblk;
    if (ncrs) *ncrs = A->pat.ncrs;
    if (vcrs) *vcrs = A->vcrs.p;
    if (vdia) *vdia = A->vdia.p;
  });
}

int dfm2h_bsr_matvec(dfm2h_bsr A, double *y, double alpha, const double *x, double beta) {
  return guard(__func__, [&] {
    need_pattern(A);
    DFM2H_REQUIRE(y && x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_REQUIRE(y != x, DFM2H_ERR_INVALID, "y and x must not alias");
    if (!A->nblk) return;
    ensure_packed(A->plan, A->pk64, A->vcrs.p, A->vdia.p, A->version, stream());
    hipLaunchKernelGGL((k_sliced_matvec<double, 3>), dim3(slice_grid(A->plan.nslices)), dim3(kPackBlock),
                       0, stream(), 0u, A->plan.nslices, view(A->plan), A->pk64.val.p,
                       A->pk64.dia.p, y, alpha, x, beta);
    check_launch("k_sliced_matvec");
  });
}

int dfm2h_bsr_set_fixed_bc(dfm2h_bsr A, const int *flag) {
  return guard(__func__, [&] {
    need_pattern(A);
    DFM2H_REQUIRE(flag, DFM2H_ERR_INVALID, "null flag");
    const size_t n = A->pat.ncrs + A->nblk;
    if (!n) return;
    hipLaunchKernelGGL(k_set_fixed_bc, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), A->nblk,
                       A->pat.ncrs, A->pat.rowof_d.p, A->pat.rowptr_d.p, A->vcrs.p, A->vdia.p,
                       flag);
    check_launch("k_set_fixed_bc");
    A->version++;
  });
}

int dfm2h_set_rhs_zero(double *b, const int *flag, size_t n, int nz) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE((b && flag) || n == 0, DFM2H_ERR_INVALID, "null vector");
    if (!n) return;
    hipLaunchKernelGGL(k_set_rhs_zero, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), b, flag, n,
                       nz);
    check_launch("k_set_rhs_zero");
  });
}

int dfm2h_xplusay(double *X, size_t n, const int *flag, double alpha, const double *Y) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE((X && flag && Y) || n == 0, DFM2H_ERR_INVALID, "null vector");
    if (!n) return;
    hipLaunchKernelGGL(k_xplusay, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), X, n, flag,
                       alpha, Y);
    check_launch("k_xplusay");
  });
}

int dfm2h_merge_linsys_solidlinear_static_tet(dfm2h_bsr A, double *vec_b, const double *myu_e,
                                              const double *lambda_e, double myu, double lambda,
                                              double rho, const double *g, const double *xyz,
                                              dfm2h_tetmesh M, const double *disp) {
  return guard(__func__, [&] {
    need_pattern(A);
    DFM2H_REQUIRE(M, DFM2H_ERR_INVALID, "null mesh handle");
    DFM2H_REQUIRE(vec_b && g && xyz, DFM2H_ERR_INVALID, "null argument");
    DFM2H_REQUIRE(M->np == A->nblk, DFM2H_ERR_INVALID, "mesh node count != matrix block rows");
    const size_t n = A->pat.ncrs + A->nblk;
    if (!n) return;
    DBuf<int> missing;
    missing.alloc(1);
    missing.zero();
    AsmArgs a;
    a.nblk = A->nblk;
    a.ncrs = A->pat.ncrs;
    a.colind = A->pat.colind_d.p;
    a.rowptr = A->pat.rowptr_d.p;
    a.rowof = A->pat.rowof_d.p;
    a.vcrs = A->vcrs.p;
    a.vdia = A->vdia.p;
    a.vec_b = vec_b;
    a.myu_e = myu_e;
    a.lambda_e = lambda_e;
    a.myu = myu;
    a.lambda = lambda;
    a.rho = rho;
    a.g0 = g[0];
    a.g1 = g[1];
    a.g2 = g[2];
    a.xyz = xyz;
    a.disp = disp;
    a.tet = reinterpret_cast<const uint4 *>(M->tet.p);
    a.elsup_ind = M->elsup_ind.p;
    a.elsup = M->elsup.p;
    a.missing = missing.p;
    hipLaunchKernelGGL(k_assemble_tet_gather, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), a);
    check_launch("k_assemble_tet_gather");
    int h_missing = 0;
    DFM2H_HIP(hipMemcpyAsync(&h_missing, missing.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
    A->version++;
    DFM2H_REQUIRE(!h_missing, DFM2H_ERR_PATTERN,
                  "an element edge has no block in the matrix pattern");
  });
}

}  // extern "C"
