/*
 * oracle.c -- CPU restatement of DelFEM2's implicit tet-FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Build with -ffp-contract=off: every expression below keeps the reference's
 * association so that results are bit-identical to the reference.
 * Citations are relative to the reference tree's include/delfem2/.
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* gridvoxel.cpp:194-271 -- nodes x-major (:210-218), 6 Kuhn tets per voxel
 * taken from the 8 corner indices (:228-267). */
void orc_mesh_tet_voxel(unsigned nx, unsigned ny, unsigned nz, double *xyz,
                        unsigned *tet) {
  const unsigned mx = nx + 1, my = ny + 1, mz = nz + 1;
  size_t ip = 0;
  for (unsigned ix = 0; ix < mx; ++ix)
    for (unsigned iy = 0; iy < my; ++iy)
      for (unsigned iz = 0; iz < mz; ++iz) {
        xyz[ip * 3 + 0] = (double)ix;
        xyz[ip * 3 + 1] = (double)iy;
        xyz[ip * 3 + 2] = (double)iz;
        ++ip;
      }
  /* corner c of voxel: bit pattern (dx,dy,dz) as in the reference corner map */
  static const unsigned cdx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  static const unsigned cdy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  static const unsigned cdz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  static const unsigned split[6][4] = {{0, 1, 2, 6}, {0, 1, 6, 5}, {0, 4, 5, 6},
                                       {0, 2, 3, 6}, {0, 3, 7, 6}, {0, 4, 6, 7}};
  size_t it = 0;
  for (unsigned ix = 0; ix < nx; ++ix)
    for (unsigned iy = 0; iy < ny; ++iy)
      for (unsigned iz = 0; iz < nz; ++iz) {
        unsigned c[8];
        for (int k = 0; k < 8; ++k)
          c[k] = (ix + cdx[k]) * (my * mz) + (iy + cdy[k]) * mz + (iz + cdz[k]);
        for (int s = 0; s < 6; ++s, ++it)
          for (int k = 0; k < 4; ++k) tet[it * 4 + k] = c[split[s][k]];
      }
}

/* ------------------------------------------------------------------------ */
/* msh_topology_uniform.cpp:19-51 (elements around a point by counting sort,
 * ascending element order), :250-299 (one ring without self, marked with a
 * per-point flag), :301-316 (driver); jagarray.cpp:72-92 (ascending sort of
 * each row -- any correct sort yields the same rows since keys are distinct). */
size_t orc_psup(const unsigned *elem, size_t nelem, unsigned nnoel, size_t np,
                unsigned *psup_ind, unsigned *psup) {
  unsigned *elsup_ind = (unsigned *)calloc(np + 1, sizeof(unsigned));
  for (size_t ie = 0; ie < nelem; ++ie)
    for (unsigned k = 0; k < nnoel; ++k) elsup_ind[elem[ie * nnoel + k] + 1] += 1;
  for (size_t ip = 0; ip < np; ++ip) elsup_ind[ip + 1] += elsup_ind[ip];
  unsigned *elsup = (unsigned *)malloc(sizeof(unsigned) * (elsup_ind[np] + 1));
  unsigned *cur = (unsigned *)malloc(sizeof(unsigned) * (np + 1));
  memcpy(cur, elsup_ind, sizeof(unsigned) * (np + 1));
  for (size_t ie = 0; ie < nelem; ++ie)
    for (unsigned k = 0; k < nnoel; ++k) elsup[cur[elem[ie * nnoel + k]]++] = (unsigned)ie;
  free(cur);

  unsigned *flg = (unsigned *)malloc(sizeof(unsigned) * (np + 1));
  for (size_t ip = 0; ip < np; ++ip) flg[ip] = UINT_MAX;
  psup_ind[0] = 0;
  for (size_t ip = 0; ip < np; ++ip) {
    unsigned cnt = 0;
    flg[ip] = (unsigned)ip;
    for (unsigned ies = elsup_ind[ip]; ies < elsup_ind[ip + 1]; ++ies) {
      const unsigned ie = elsup[ies];
      for (unsigned k = 0; k < nnoel; ++k) {
        const unsigned jp = elem[ie * nnoel + k];
        if (flg[jp] != ip) {
          flg[jp] = (unsigned)ip;
          if (psup) psup[psup_ind[ip] + cnt] = jp;
          ++cnt;
        }
      }
    }
    psup_ind[ip + 1] = psup_ind[ip] + cnt;
    if (psup) { /* insertion sort of the row, ascending */
      unsigned *row = psup + psup_ind[ip];
      for (unsigned a = 1; a < cnt; ++a) {
        const unsigned v = row[a];
        unsigned b = a;
        while (b > 0 && row[b - 1] > v) { row[b] = row[b - 1]; --b; }
        row[b] = v;
      }
    }
  }
  free(flg);
  free(elsup);
  free(elsup_ind);
  return psup_ind[np];
}

/* ------------------------------------------------------------------------ */
/* femutil.cpp:202-212: signed volume, scalar triple product times 1/6. */
double orc_tet_volume(const double *v1, const double *v2, const double *v3,
                      const double *v4) {
  const double one_sixth = 1.0 / 6.0;
  return ((v2[0] - v1[0]) * ((v3[1] - v1[1]) * (v4[2] - v1[2]) - (v4[1] - v1[1]) * (v3[2] - v1[2]))
          - (v2[1] - v1[1]) * ((v3[0] - v1[0]) * (v4[2] - v1[2]) - (v4[0] - v1[0]) * (v3[2] - v1[2]))
          + (v2[2] - v1[2]) * ((v3[0] - v1[0]) * (v4[1] - v1[1]) - (v4[0] - v1[0]) * (v3[1] - v1[1]))) *
         one_sixth;
}

/* femutil.cpp:215-258: gradient of the barycentric coordinate of node n is
 * the (signed) normal of the opposite face over 6*vol.  Node n uses the other
 * three nodes in cyclic order (n+1, n+2, n+3) and the sign alternates
 * (-,+,-,+) with n; the component signs are (s,-s,s). */
void orc_tet_dldx(double dldx[4][3], const double *p0, const double *p1,
                  const double *p2, const double *p3) {
  const double vol = orc_tet_volume(p0, p1, p2, p3);
  const double dtmp1 = 1.0 / (vol * 6.0);
  const double *p[4] = {p0, p1, p2, p3};
  for (int n = 0; n < 4; ++n) {
    const double *qa = p[(n + 1) & 3], *qb = p[(n + 2) & 3], *qc = p[(n + 3) & 3];
    const double s = (n & 1) ? dtmp1 : -dtmp1;
    dldx[n][0] = s * ((qb[1] - qa[1]) * (qc[2] - qa[2]) - (qc[1] - qa[1]) * (qb[2] - qa[2]));
    dldx[n][1] = -s * ((qb[0] - qa[0]) * (qc[2] - qa[2]) - (qc[0] - qa[0]) * (qb[2] - qa[2]));
    dldx[n][2] = s * ((qb[0] - qa[0]) * (qc[1] - qa[1]) - (qc[0] - qa[0]) * (qb[1] - qa[1]));
  }
}

/* fem_solidlinear.cpp:409-435 with :55-79 and femutil.h:252-268.
 * emat is zeroed here (the driver zero-clears it, fem_solidlinear.h:223) and
 * then accumulated additively exactly as the reference does. */
void orc_emat_tet(double emat[4][4][3][3], double eres[4][3], double myu,
                  double lambda, const double P[4][3], const double disp[4][3]) {
  const double vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
  double d[4][3];
  orc_tet_dldx(d, P[0], P[1], P[2], P[3]);
  const double w = vol;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double v = 0.0;
          v += w * (lambda * d[i][a] * d[j][b] + myu * d[j][a] * d[i][b]);
          emat[i][j][a][b] = v;
        }
      const double dtmp1 = d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2];
      emat[i][j][0][0] += w * myu * dtmp1;
      emat[i][j][1][1] += w * myu * dtmp1;
      emat[i][j][2][2] += w * myu * dtmp1;
    }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int a = 0; a < 3; ++a)
        eres[i][a] += -1. * (emat[i][j][a][0] * disp[j][0] + emat[i][j][a][1] * disp[j][1] +
                             emat[i][j][a][2] * disp[j][2]);
}

/* fem_solidlinear.h:192-237 driver, per-element Lame pair; Merge<4,4,3,3>
 * ls_block_sparse.h:260-310 (diagonal blocks to vdia, others to the slot of
 * column j in row i). */
int orc_assemble_tet_static(unsigned nblk, const unsigned *colind,
                            const unsigned *rowptr, double *vcrs, double *vdia,
                            double *vec_b, const double *myu_e,
                            const double *lambda_e, double rho, const double *g,
                            const double *xyz, const unsigned *tet, size_t ntet,
                            const double *disp) {
  unsigned *buf = (unsigned *)malloc(sizeof(unsigned) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) buf[i] = UINT_MAX;
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *ip = tet + iel * 4;
    double P[4][3], U[4][3];
    for (int k = 0; k < 4; ++k)
      for (int a = 0; a < 3; ++a) {
        P[k][a] = xyz[ip[k] * 3 + a];
        U[k][a] = disp ? disp[ip[k] * 3 + a] : 0.0;
      }
    double eres[4][3];
    {
      const double vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
      for (int k = 0; k < 4; ++k) {
        eres[k][0] = vol * rho * g[0] * 0.25;
        eres[k][1] = vol * rho * g[1] * 0.25;
        eres[k][2] = vol * rho * g[2] * 0.25;
      }
    }
    double emat[4][4][3][3];
    orc_emat_tet(emat, eres, myu_e[iel], lambda_e[iel], (const double(*)[3])P,
                 (const double(*)[3])U);
    for (int k = 0; k < 4; ++k) {
      vec_b[ip[k] * 3 + 0] += eres[k][0];
      vec_b[ip[k] * 3 + 1] += eres[k][1];
      vec_b[ip[k] * 3 + 2] += eres[k][2];
    }
    for (int ir = 0; ir < 4; ++ir) {
      const unsigned i = ip[ir];
      for (unsigned s = colind[i]; s < colind[i + 1]; ++s) buf[rowptr[s]] = s;
      for (int jc = 0; jc < 4; ++jc) {
        const unsigned j = ip[jc];
        double *out;
        if (i == j) {
          out = vdia + (size_t)i * 9;
        } else {
          if (buf[j] == UINT_MAX) { status = 1; continue; }
          out = vcrs + (size_t)buf[j] * 9;
        }
        const double *in = &emat[ir][jc][0][0];
        for (int k = 0; k < 9; ++k) out[k] += in[k];
      }
      for (unsigned s = colind[i]; s < colind[i + 1]; ++s) buf[rowptr[s]] = UINT_MAX;
    }
  }
  free(buf);
  return status;
}

/* ls_block_sparse.cpp:397-415 (diagonal), :427-444 (row), :456-472 (column). */
void orc_set_fixed_bc(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, double *vcrs, double *vdia,
                      const int *bc_flag) {
  for (unsigned ib = 0; ib < nblk; ++ib)
    for (unsigned il = 0; il < 3; ++il) {
      if (bc_flag[ib * 3 + il] == 0) continue;
      for (unsigned jl = 0; jl < 3; ++jl) {
        vdia[ib * 9 + il * 3 + jl] = 0.0;
        vdia[ib * 9 + jl * 3 + il] = 0.0;
      }
      vdia[ib * 9 + il * 3 + il] = 1.0;
    }
  for (unsigned ib = 0; ib < nblk; ++ib)
    for (unsigned s = colind[ib]; s < colind[ib + 1]; ++s)
      for (unsigned il = 0; il < 3; ++il) {
        if (bc_flag[ib * 3 + il] == 0) continue;
        for (unsigned jl = 0; jl < 3; ++jl) vcrs[(size_t)s * 9 + il * 3 + jl] = 0.0;
      }
  const unsigned ncrs = colind[nblk];
  for (unsigned s = 0; s < ncrs; ++s) {
    const unsigned jb = rowptr[s];
    for (unsigned jl = 0; jl < 3; ++jl) {
      if (bc_flag[jb * 3 + jl] == 0) continue;
      for (unsigned il = 0; il < 3; ++il) vcrs[(size_t)s * 9 + il * 3 + jl] = 0.0;
    }
  }
}

/* vecxitrsol.cpp:286-296 */
void orc_set_rhs_zero(double *b, const int *flag, size_t n, int iflag_nonzero) {
  for (size_t i = 0; i < n; ++i) {
    if (flag[i] == iflag_nonzero) continue;
    b[i] = 0;
  }
}

/* ls_block_sparse.cpp:215-222 (y *= beta first) and :103-133 (off-diagonal
 * blocks in stored order, then the diagonal block). */
void orc_matvec(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                const double *vcrs, const double *vdia, double *y,
                double alpha, const double *x, double beta) {
  const size_t n = (size_t)nblk * 3;
  for (size_t i = 0; i < n; ++i) y[i] *= beta;
  for (unsigned ib = 0; ib < nblk; ++ib) {
    const size_t i0 = (size_t)ib * 3;
    for (unsigned s = colind[ib]; s < colind[ib + 1]; ++s) {
      const size_t j0 = (size_t)rowptr[s] * 3;
      const double *v = vcrs + (size_t)s * 9;
      y[i0 + 0] += alpha * (v[0] * x[j0 + 0] + v[1] * x[j0 + 1] + v[2] * x[j0 + 2]);
      y[i0 + 1] += alpha * (v[3] * x[j0 + 0] + v[4] * x[j0 + 1] + v[5] * x[j0 + 2]);
      y[i0 + 2] += alpha * (v[6] * x[j0 + 0] + v[7] * x[j0 + 1] + v[8] * x[j0 + 2]);
    }
    const double *v = vdia + (size_t)ib * 9;
    y[i0 + 0] += alpha * (v[0] * x[i0 + 0] + v[1] * x[i0 + 1] + v[2] * x[i0 + 2]);
    y[i0 + 1] += alpha * (v[3] * x[i0 + 0] + v[4] * x[i0 + 1] + v[5] * x[i0 + 2]);
    y[i0 + 2] += alpha * (v[6] * x[i0 + 0] + v[7] * x[i0 + 1] + v[8] * x[i0 + 2]);
  }
}

/* ------------------------------------------------------------------------ */
/* ls_ilu_block_sparse.cpp:1057-1090 */
static int cmp_uint(const void *a, const void *b) {
  const unsigned x = *(const unsigned *)a, y = *(const unsigned *)b;
  return (x > y) - (x < y);
}
void orc_ilu_set_pattern0(unsigned nblk, const unsigned *colind,
                          const unsigned *rowptr, unsigned *rowptr_out,
                          unsigned *diaind) {
  memcpy(rowptr_out, rowptr, sizeof(unsigned) * colind[nblk]);
  for (unsigned ib = 0; ib < nblk; ++ib)
    qsort(rowptr_out + colind[ib], colind[ib + 1] - colind[ib], sizeof(unsigned), cmp_uint);
  for (unsigned ib = 0; ib < nblk; ++ib) {
    diaind[ib] = colind[ib + 1];
    for (unsigned s = colind[ib]; s < colind[ib + 1]; ++s)
      if (rowptr_out[s] > ib) { diaind[ib] = s; break; }
  }
}

/* ls_ilu_block_sparse.cpp:1009-1047 */
void orc_ilu_copy_value(unsigned nblk, const unsigned *ilu_colind,
                        const unsigned *ilu_rowptr, double *ilu_vcrs,
                        double *ilu_vdia, const unsigned *a_colind,
                        const unsigned *a_rowptr, const double *a_vcrs,
                        const double *a_vdia) {
  int *row2crs = (int *)malloc(sizeof(int) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) row2crs[i] = -1;
  const size_t n = (size_t)ilu_colind[nblk] * 9;
  for (size_t i = 0; i < n; ++i) ilu_vcrs[i] = 0.0;
  for (unsigned ib = 0; ib < nblk; ++ib) {
    for (unsigned s = ilu_colind[ib]; s < ilu_colind[ib + 1]; ++s) row2crs[ilu_rowptr[s]] = (int)s;
    for (unsigned s = a_colind[ib]; s < a_colind[ib + 1]; ++s) {
      const int d = row2crs[a_rowptr[s]];
      if (d == -1) continue;
      for (int k = 0; k < 9; ++k) ilu_vcrs[(size_t)d * 9 + k] = a_vcrs[(size_t)s * 9 + k];
    }
    for (unsigned s = ilu_colind[ib]; s < ilu_colind[ib + 1]; ++s) row2crs[ilu_rowptr[s]] = -1;
  }
  for (size_t i = 0; i < (size_t)nblk * 9; ++i) ilu_vdia[i] = a_vdia[i];
  free(row2crs);
}

/* ls_ilu_block_sparse.cpp:84-103 CalcInvMat3 */
static double det3(const double *a) {
  return +a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5]
         - a[0] * a[7] * a[5] - a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
}
static void inv3_inplace(double *a) {
  const double inv_det = 1.0 / det3(a);
  double t[9];
  for (int i = 0; i < 9; ++i) t[i] = a[i];
  a[0] = inv_det * (t[4] * t[8] - t[5] * t[7]);
  a[1] = inv_det * (t[2] * t[7] - t[1] * t[8]);
  a[2] = inv_det * (t[1] * t[5] - t[2] * t[4]);
  a[3] = inv_det * (t[5] * t[6] - t[3] * t[8]);
  a[4] = inv_det * (t[0] * t[8] - t[2] * t[6]);
  a[5] = inv_det * (t[2] * t[3] - t[0] * t[5]);
  a[6] = inv_det * (t[3] * t[7] - t[4] * t[6]);
  a[7] = inv_det * (t[1] * t[6] - t[0] * t[7]);
  a[8] = inv_det * (t[0] * t[4] - t[1] * t[3]);
}

/* ls_ilu_block_sparse.cpp:152-156, 286-356: row-sequential IKJ block ILU(0). */
int orc_ilu_decompose(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, const unsigned *diaind,
                      double *vcrs, double *vdia, int *n_singular) {
  int nsing = 0;
  int *row2crs = (int *)malloc(sizeof(int) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) row2crs[i] = -1;
  int ok = 1;
  for (unsigned ib = 0; ib < nblk && ok; ++ib) {
    for (unsigned s = colind[ib]; s < colind[ib + 1]; ++s) row2crs[rowptr[s]] = (int)s;
    for (unsigned ik = colind[ib]; ik < diaind[ib]; ++ik) {
      const unsigned kb = rowptr[ik];
      const double *vik = vcrs + (size_t)ik * 9;
      for (unsigned kj = diaind[kb]; kj < colind[kb + 1]; ++kj) {
        const unsigned jb = rowptr[kj];
        const double *vkj = vcrs + (size_t)kj * 9;
        double *vij;
        if (jb != ib) {
          const int ij = row2crs[jb];
          if (ij == -1) continue;
          vij = vcrs + (size_t)ij * 9;
        } else {
          vij = vdia + (size_t)ib * 9;
        }
        for (int i = 0; i < 3; ++i) {
          vij[i * 3 + 0] -= vik[i * 3 + 0] * vkj[0] + vik[i * 3 + 1] * vkj[3] + vik[i * 3 + 2] * vkj[6];
          vij[i * 3 + 1] -= vik[i * 3 + 0] * vkj[1] + vik[i * 3 + 1] * vkj[4] + vik[i * 3 + 2] * vkj[7];
          vij[i * 3 + 2] -= vik[i * 3 + 0] * vkj[2] + vik[i * 3 + 1] * vkj[5] + vik[i * 3 + 2] * vkj[8];
        }
      }
    }
    {
      double *vii = vdia + (size_t)ib * 9;
      const double det = det3(vii);
      if (fabs(det) > 1.0e-30) {
        inv3_inplace(vii);
      } else {
        ++nsing;
        if (nsing > 10) { ok = 0; break; }
      }
    }
    {
      const double *vii = vdia + (size_t)ib * 9;
      for (unsigned s = diaind[ib]; s < colind[ib + 1]; ++s) {
        double *vij = vcrs + (size_t)s * 9;
        double t[9];
        for (int k = 0; k < 9; ++k) t[k] = vij[k];
        for (int i = 0; i < 3; ++i) {
          vij[i * 3 + 0] = vii[i * 3 + 0] * t[0] + vii[i * 3 + 1] * t[3] + vii[i * 3 + 2] * t[6];
          vij[i * 3 + 1] = vii[i * 3 + 0] * t[1] + vii[i * 3 + 1] * t[4] + vii[i * 3 + 2] * t[7];
          vij[i * 3 + 2] = vii[i * 3 + 0] * t[2] + vii[i * 3 + 1] * t[5] + vii[i * 3 + 2] * t[8];
        }
      }
    }
    for (unsigned s = colind[ib]; s < colind[ib + 1]; ++s) row2crs[rowptr[s]] = -1;
  }
  free(row2crs);
  if (n_singular) *n_singular = nsing;
  return ok;
}

/* ls_ilu_block_sparse.cpp:544-573 */
void orc_ilu_forward(unsigned nblk, const unsigned *colind,
                     const unsigned *rowptr, const unsigned *diaind,
                     const double *vcrs, const double *vdia, double *vec) {
  for (unsigned ib = 0; ib < nblk; ++ib) {
    double t0 = vec[ib * 3 + 0], t1 = vec[ib * 3 + 1], t2 = vec[ib * 3 + 2];
    for (unsigned s = colind[ib]; s < diaind[ib]; ++s) {
      const size_t j0 = (size_t)rowptr[s] * 3;
      const double *v = vcrs + (size_t)s * 9;
      const double x0 = vec[j0 + 0], x1 = vec[j0 + 1], x2 = vec[j0 + 2];
      t0 -= v[0] * x0 + v[1] * x1 + v[2] * x2;
      t1 -= v[3] * x0 + v[4] * x1 + v[5] * x2;
      t2 -= v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    const double *d = vdia + (size_t)ib * 9;
    vec[ib * 3 + 0] = d[0] * t0 + d[1] * t1 + d[2] * t2;
    vec[ib * 3 + 1] = d[3] * t0 + d[4] * t1 + d[5] * t2;
    vec[ib * 3 + 2] = d[6] * t0 + d[7] * t1 + d[8] * t2;
  }
}

/* ls_ilu_block_sparse.cpp:721-749 */
void orc_ilu_backward(unsigned nblk, const unsigned *colind,
                      const unsigned *rowptr, const unsigned *diaind,
                      const double *vcrs, double *vec) {
  for (unsigned ib = nblk - 1; ib != UINT_MAX; --ib) {
    double t0 = vec[ib * 3 + 0], t1 = vec[ib * 3 + 1], t2 = vec[ib * 3 + 2];
    for (unsigned s = diaind[ib]; s < colind[ib + 1]; ++s) {
      const size_t j0 = (size_t)rowptr[s] * 3;
      const double *v = vcrs + (size_t)s * 9;
      const double x0 = vec[j0 + 0], x1 = vec[j0 + 1], x2 = vec[j0 + 2];
      t0 -= v[0] * x0 + v[1] * x1 + v[2] * x2;
      t1 -= v[3] * x0 + v[4] * x1 + v[5] * x2;
      t2 -= v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    vec[ib * 3 + 0] = t0;
    vec[ib * 3 + 1] = t1;
    vec[ib * 3 + 2] = t2;
  }
}

/* ------------------------------------------------------------------------ */
/* view_vectorx.h:32-39 */
double orc_dot(const double *x, const double *y, size_t n) {
  double d = 0;
  for (size_t i = 0; i < n; ++i) d += x[i] * y[i];
  return d;
}

/* lsitrsol.h:80-148 with view_vectorx.h:61-109 (AddScaledVec y += a*x,
 * ScaleAndAddVec y = b*y + x, AddMatVec -> MatVec(y, 1, p, 0)). */
size_t orc_solve_pcg(double *r, double *x, double tol, unsigned max_nitr,
                     unsigned nblk, const unsigned *a_colind,
                     const unsigned *a_rowptr, const double *a_vcrs,
                     const double *a_vdia, const unsigned *l_colind,
                     const unsigned *l_rowptr, const unsigned *l_diaind,
                     const double *l_vcrs, const double *l_vdia, double *hist,
                     size_t hist_cap) {
  const size_t n = (size_t)nblk * 3;
  size_t nh = 0;
#define PUSH(v) do { if (nh < hist_cap) hist[nh] = (v); ++nh; } while (0)
  for (size_t i = 0; i < n; ++i) x[i] = 0;
  const double sq0 = orc_dot(r, r, n);
  PUSH(sqrt(sq0));
  if (sq0 < 1.0e-30) return nh;
  const double inv_sq0 = 1.0 / sq0;
  double *Pr = (double *)malloc(sizeof(double) * (n + 1));
  double *p = (double *)malloc(sizeof(double) * (n + 1));
  memcpy(Pr, r, sizeof(double) * n);
  orc_ilu_forward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, l_vdia, Pr);
  orc_ilu_backward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, Pr);
  memcpy(p, Pr, sizeof(double) * n);
  double rPr = orc_dot(r, Pr, n);
  int converged = 0;
  for (unsigned it = 0; it < max_nitr; ++it) {
    double *Ap = Pr;
    orc_matvec(nblk, a_colind, a_rowptr, a_vcrs, a_vdia, Ap, 1.0, p, 0.0);
    const double pAp = orc_dot(p, Ap, n);
    const double alpha = rPr / pAp;
    for (size_t i = 0; i < n; ++i) r[i] += -alpha * Ap[i];
    for (size_t i = 0; i < n; ++i) x[i] += +alpha * p[i];
    {
      const double sq = orc_dot(r, r, n);
      PUSH(sqrt(sq));
      const double ratio = sqrt(sq * inv_sq0);
      if (ratio < tol) { converged = 1; break; }
    }
    memcpy(Pr, r, sizeof(double) * n);
    orc_ilu_forward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, l_vdia, Pr);
    orc_ilu_backward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, Pr);
    const double rPr1 = orc_dot(r, Pr, n);
    const double beta = rPr1 / rPr;
    rPr = rPr1;
    for (size_t i = 0; i < n; ++i) p[i] = beta * p[i] + Pr[i];
  }
  if (!converged) PUSH(sqrt(orc_dot(r, r, n)));
#undef PUSH
  free(Pr);
  free(p);
  return nh;
}

/* vecxitrsol.cpp:226-237 */
void orc_xplusay(double *X, size_t n, const int *flag, double alpha,
                 const double *Y) {
  for (size_t i = 0; i < n; ++i) {
    if (flag[i] != 0) continue;
    X[i] += alpha * Y[i];
  }
}

/* Adjoint contraction: K_e is linear in (myu, lambda), so dK_e/dE_e is the
 * element matrix at the E=1 Lame pair.  g = -lam_e^T K_e(E=1) u_e, summed
 * over the 16 blocks in (i, j, a, b) order. */
void orc_adjoint_grad_tet(double *grad, const double *mu1_e,
                          const double *lambda1_e, const double *xyz,
                          const unsigned *tet, size_t ntet, const double *u,
                          const double *lam) {
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *ip = tet + iel * 4;
    double P[4][3], Z[4][3], emat[4][4][3][3], eres[4][3];
    for (int k = 0; k < 4; ++k)
      for (int a = 0; a < 3; ++a) {
        P[k][a] = xyz[ip[k] * 3 + a];
        Z[k][a] = 0.0;
        eres[k][a] = 0.0;
      }
    orc_emat_tet(emat, eres, mu1_e[iel], lambda1_e[iel], (const double(*)[3])P,
                 (const double(*)[3])Z);
    double s = 0.0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            s += lam[ip[i] * 3 + a] * emat[i][j][a][b] * u[ip[j] * 3 + b];
    grad[iel] = -s;
  }
}

/* ------------------------------------------------------------------------ */
/* ls_ilu_block_sparse.cpp:849-999 Initialize_ILUk, symbolic phase for
 * lev_fill >= 1 (or -1: every fill).  Row by row: the row's entries start as
 * A's columns (sorted, level 0); the entries k < i are visited in ascending
 * order, including fills added meanwhile; a k whose level + 1 exceeds
 * lev_fill is skipped; for every entry j of row k's upper part (levels
 * likewise limited), j != i, not yet in row i: add (j, max(lev_ik, lev_kj) +
 * 1).  An entry already present keeps its level (the reference's rule).
 * Output: ILU colind[nblk+1]; rowptr / diaind (first column > i) filled when
 * cap >= the returned entry count.  Empty rows are allowed here. */
typedef struct { unsigned col, lev; } orc_cl;

static int orc_cl_cmp(const void *a, const void *b) {
  const unsigned x = ((const orc_cl *)a)->col, y = ((const orc_cl *)b)->col;
  return x < y ? -1 : x > y;
}

size_t orc_ilu_pattern_k(unsigned nblk, const unsigned *a_colind, const unsigned *a_rowptr,
                         int lev_fill, unsigned *colind, unsigned *rowptr, unsigned *diaind,
                         size_t cap) {
  size_t ncap = (size_t)a_colind[nblk] * 2 + 16, nent = 0;
  orc_cl *ent = (orc_cl *)malloc(sizeof(orc_cl) * ncap);
  unsigned *dia = (unsigned *)malloc(sizeof(unsigned) * (nblk + 1));
  size_t rcap = 64;
  orc_cl *row = (orc_cl *)malloc(sizeof(orc_cl) * rcap);
  colind[0] = 0;
  for (unsigned i = 0; i < nblk; ++i) {
    size_t n = 0;
    for (unsigned s = a_colind[i]; s < a_colind[i + 1]; ++s) {
      if (n == rcap) { rcap *= 2; row = (orc_cl *)realloc(row, sizeof(orc_cl) * rcap); }
      row[n].col = a_rowptr[s];
      row[n].lev = 0;
      ++n;
    }
    qsort(row, n, sizeof(orc_cl), orc_cl_cmp);
    for (size_t t = 0; t < n; ++t) {
      const unsigned k = row[t].col, lik = row[t].lev;
      if (lev_fill != -1 && (int)lik + 1 > lev_fill) continue;
      if (k >= i) break;
      for (size_t kj = dia[k]; kj < colind[k + 1]; ++kj) {
        const unsigned lkj = ent[kj].lev, j = ent[kj].col;
        if (lev_fill != -1 && (int)lkj + 1 > lev_fill) continue;
        if (j == i) continue;
        /* position of j among row[t+1 .. n) (j > k) */
        size_t lo = t + 1, hi = n;
        while (lo < hi) {
          const size_t mid = (lo + hi) / 2;
          if (row[mid].col < j) lo = mid + 1; else hi = mid;
        }
        if (lo < n && row[lo].col == j) continue;
        if (n == rcap) { rcap *= 2; row = (orc_cl *)realloc(row, sizeof(orc_cl) * rcap); }
        memmove(row + lo + 1, row + lo, sizeof(orc_cl) * (n - lo));
        row[lo].col = j;
        row[lo].lev = (lik > lkj ? lik : lkj) + 1;
        ++n;
      }
    }
    if (nent + n > ncap) {
      while (nent + n > ncap) ncap *= 2;
      ent = (orc_cl *)realloc(ent, sizeof(orc_cl) * ncap);
    }
    memcpy(ent + nent, row, sizeof(orc_cl) * n);
    colind[i + 1] = (unsigned)(nent + n);
    dia[i] = colind[i + 1];
    for (size_t t = 0; t < n; ++t)
      if (row[t].col > i) { dia[i] = (unsigned)(nent + t); break; }
    nent += n;
  }
  if (rowptr && cap >= nent) {
    for (size_t t = 0; t < nent; ++t) rowptr[t] = ent[t].col;
    if (diaind) memcpy(diaind, dia, sizeof(unsigned) * nblk);
  }
  free(ent);
  free(dia);
  free(row);
  return nent;
}

/* ------------------------------------------------------------------------ */
/* Time-stepping drivers. */

/* fem_solidlinear.h:129-153 SetEmat_LinearSolid3<4>: assigned, so an exact
 * zero keeps its sign (the additive form on a zeroed emat would turn -0 into
 * +0). */
static void set_emat_ls3(double e[4][4][3][3], double lambda, double myu, const double d[4][3],
                         double w) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          e[i][j][a][b] = w * (lambda * d[i][a] * d[j][b] + myu * d[j][a] * d[i][b]);
      const double dtmp1 = d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2];
      e[i][j][0][0] += w * myu * dtmp1;
      e[i][j][1][1] += w * myu * dtmp1;
      e[i][j][2][2] += w * myu * dtmp1;
    }
}

/* femutil.cpp:260-280 SetEmatConsistentMassTet */
static void set_emat_mass_tet(double m[4][4][3][3], double w0) {
  const double dtmp1 = w0 * 0.05;
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) m[i][j][a][b] = a == b ? dtmp1 : 0.0;
    m[i][i][0][0] += dtmp1;
    m[i][i][1][1] += dtmp1;
    m[i][i][2][2] += dtmp1;
  }
}

/* femutil.h:252-268 AddEmatEvecScale3<4> */
static void add_emat_evec_scale3(double R[4][3], const double e[4][4][3][3],
                                 const double v[4][3], double scale) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int a = 0; a < 3; ++a)
        R[i][a] += scale * (e[i][j][a][0] * v[j][0] + e[i][j][a][1] * v[j][1] +
                            e[i][j][a][2] * v[j][2]);
}

/* ls_block_sparse.h:260-310 Merge<4,4,3,3>: row i's column -> slot map in buf
 * (all UINT_MAX on entry and exit); returns 1 on a missing block. */
static int merge_tet(const unsigned *ip, const double e[4][4][3][3], const unsigned *colind,
                     const unsigned *rowptr, double *vcrs, double *vdia, unsigned *buf) {
  int status = 0;
  for (int ir = 0; ir < 4; ++ir) {
    const unsigned i = ip[ir];
    for (unsigned s = colind[i]; s < colind[i + 1]; ++s) buf[rowptr[s]] = s;
    for (int jc = 0; jc < 4; ++jc) {
      const unsigned j = ip[jc];
      double *out;
      if (i == j) {
        out = vdia + (size_t)i * 9;
      } else {
        if (buf[j] == UINT_MAX) { status = 1; continue; }
        out = vcrs + (size_t)buf[j] * 9;
      }
      const double *in = &e[ir][jc][0][0];
      for (int k = 0; k < 9; ++k) out[k] += in[k];
    }
    for (unsigned s = colind[i]; s < colind[i + 1]; ++s) buf[rowptr[s]] = UINT_MAX;
  }
  return status;
}

static void fetch43(double out[4][3], const unsigned *ip, const double *v) {
  for (int k = 0; k < 4; ++k)
    for (int a = 0; a < 3; ++a) out[k][a] = v[ip[k] * 3 + a];
}

int orc_assemble_tet_newmark(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                             double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                             const double *lambda_e, double rho, const double *g, double dt,
                             double beta_newmark, const double *xyz, const unsigned *tet,
                             size_t ntet, const double *disp, const double *velo,
                             const double *acc) {
  unsigned *buf = (unsigned *)malloc(sizeof(unsigned) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) buf[i] = UINT_MAX;
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *ip = tet + iel * 4;
    double P[4][3], U[4][3], V[4][3], A[4][3];
    fetch43(P, ip, xyz);
    fetch43(U, ip, disp);
    fetch43(V, ip, velo);
    fetch43(A, ip, acc);
    /* fem_solidlinear.cpp:437-490 */
    const double vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
    double d[4][3];
    orc_tet_dldx(d, P[0], P[1], P[2], P[3]);
    double eK[4][4][3][3], eM[4][4][3][3], emat[4][4][3][3], eres[4][3];
    set_emat_ls3(eK, lambda_e[iel], myu_e[iel], (const double(*)[3])d, vol);
    set_emat_mass_tet(eM, rho * vol);
    const double dtmp1 = beta_newmark * dt * dt;
    for (int k = 0; k < 144; ++k)
      (&emat[0][0][0][0])[k] = (&eM[0][0][0][0])[k] + dtmp1 * (&eK[0][0][0][0])[k];
    for (int k = 0; k < 4; ++k) {
      eres[k][0] = vol * rho * g[0] * 0.25;
      eres[k][1] = vol * rho * g[1] * 0.25;
      eres[k][2] = vol * rho * g[2] * 0.25;
    }
    add_emat_evec_scale3(eres, (const double(*)[4][3][3])eK, (const double(*)[3])U, -1.);
    add_emat_evec_scale3(eres, (const double(*)[4][3][3])eM, (const double(*)[3])A, -1.);
    add_emat_evec_scale3(eres, (const double(*)[4][3][3])eK, (const double(*)[3])V, -dt);
    add_emat_evec_scale3(eres, (const double(*)[4][3][3])eK, (const double(*)[3])A,
                         -0.5 * dt * dt);
    for (int k = 0; k < 4; ++k) {
      vec_b[ip[k] * 3 + 0] += eres[k][0];
      vec_b[ip[k] * 3 + 1] += eres[k][1];
      vec_b[ip[k] * 3 + 2] += eres[k][2];
    }
    status |= merge_tet(ip, (const double(*)[4][3][3])emat, colind, rowptr, vcrs, vdia, buf);
  }
  free(buf);
  return status;
}

int orc_assemble_tet_beuler(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                            double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                            const double *lambda_e, double rho, const double *g, double dt,
                            const double *xyz, const unsigned *tet, size_t ntet,
                            const double *disp, const double *velo) {
  unsigned *buf = (unsigned *)malloc(sizeof(unsigned) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) buf[i] = UINT_MAX;
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *ip = tet + iel * 4;
    double P[4][3], u[4][3], v[4][3];
    fetch43(P, ip, xyz);
    const double vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
    double d[4][3], emat[4][4][3][3], eres[4][3];
    orc_tet_dldx(d, P[0], P[1], P[2], P[3]);
    set_emat_ls3(emat, lambda_e[iel], myu_e[iel], (const double(*)[3])d, vol);
    fetch43(u, ip, disp);
    fetch43(v, ip, velo);
    for (int ino = 0; ino < 4; ++ino)
      for (int a = 0; a < 3; ++a) {
        eres[ino][a] = vol * rho * g[a] * 0.25;
        for (int jno = 0; jno < 4; ++jno) {
          eres[ino][a] -= emat[ino][jno][a][0] * (u[jno][0] + dt * v[jno][0]);
          eres[ino][a] -= emat[ino][jno][a][1] * (u[jno][1] + dt * v[jno][1]);
          eres[ino][a] -= emat[ino][jno][a][2] * (u[jno][2] + dt * v[jno][2]);
        }
      }
    for (int ino = 0; ino < 4; ++ino) {
      emat[ino][ino][0][0] += rho * vol * 0.25 / (dt * dt);
      emat[ino][ino][1][1] += rho * vol * 0.25 / (dt * dt);
      emat[ino][ino][2][2] += rho * vol * 0.25 / (dt * dt);
    }
    for (int k = 0; k < 4; ++k) {
      vec_b[ip[k] * 3 + 0] += eres[k][0] / dt;
      vec_b[ip[k] * 3 + 1] += eres[k][1] / dt;
      vec_b[ip[k] * 3 + 2] += eres[k][2] / dt;
    }
    status |= merge_tet(ip, (const double(*)[4][3][3])emat, colind, rowptr, vcrs, vdia, buf);
  }
  free(buf);
  return status;
}

void orc_xplusaybz(double *X, size_t n, const int *flag, double alpha, const double *Y,
                   double beta, const double *Z) {
  for (size_t i = 0; i < n; ++i) {
    if (flag[i] != 0) continue;
    X[i] += alpha * Y[i] + beta * Z[i];
  }
}

void orc_xplusaybzcw(double *X, size_t n, const int *flag, double alpha, const double *Y,
                     double beta, const double *Z, double gamma, const double *W) {
  for (size_t i = 0; i < n; ++i) {
    if (flag[i] != 0) continue;
    X[i] += alpha * Y[i] + beta * Z[i] + gamma * W[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Stiffness warping (603_Fem3DStiffwarp) and its solver. */

/* vecxitrsol.h:350-437 Solve_PBiCGStab with the ILU preconditioner; DotX /
 * Dot sequential (vecxitrsol.cpp:20-28, 61-68), AXPY y += a x
 * (vecxitrsol.cpp:153-159).  Returns the history length. */
size_t orc_solve_pbicgstab(double *r, double *x, double tol, unsigned max_niter,
                           unsigned nblk, const unsigned *a_colind, const unsigned *a_rowptr,
                           const double *a_vcrs, const double *a_vdia,
                           const unsigned *l_colind, const unsigned *l_rowptr,
                           const unsigned *l_diaind, const double *l_vcrs,
                           const double *l_vdia, double *hist, size_t hist_cap) {
  const size_t n = (size_t)nblk * 3;
  size_t nh = 0;
#define PUSH(v) do { if (nh < hist_cap) hist[nh] = (v); ++nh; } while (0)
  for (size_t i = 0; i < n; ++i) x[i] = 0.0;
  const double sq_ini = orc_dot(r, r, n);
  if (sq_ini < 1.0e-60) {
    PUSH(sqrt(sq_ini));
    return nh;
  }
  const double sq_inv_ini = 1.0 / sq_ini;
  double *s = (double *)calloc(n + 1, sizeof(double));
  double *Ms = (double *)calloc(n + 1, sizeof(double));
  double *AMs = (double *)calloc(n + 1, sizeof(double));
  double *Mp = (double *)calloc(n + 1, sizeof(double));
  double *AMp = (double *)calloc(n + 1, sizeof(double));
  double *r0 = (double *)calloc(n + 1, sizeof(double));
  double *p = (double *)calloc(n + 1, sizeof(double));
  memcpy(r0, r, sizeof(double) * n);
  memcpy(p, r, sizeof(double) * n);
  for (unsigned it = 1; it < max_niter; ++it) {
    memcpy(Mp, p, sizeof(double) * n);
    orc_ilu_forward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, l_vdia, Mp);
    orc_ilu_backward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, Mp);
    const double r_r2 = orc_dot(r, r0, n);
    orc_matvec(nblk, a_colind, a_rowptr, a_vcrs, a_vdia, AMp, 1.0, Mp, 0.0);
    const double alpha = r_r2 / orc_dot(AMp, r0, n);
    memcpy(s, r, sizeof(double) * n);
    for (size_t i = 0; i < n; ++i) s[i] += -alpha * AMp[i];
    memcpy(Ms, s, sizeof(double) * n);
    orc_ilu_forward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, l_vdia, Ms);
    orc_ilu_backward(nblk, l_colind, l_rowptr, l_diaind, l_vcrs, Ms);
    orc_matvec(nblk, a_colind, a_rowptr, a_vcrs, a_vdia, AMs, 1.0, Ms, 0.0);
    const double denominator = orc_dot(AMs, AMs, n);
    const double numerator = orc_dot(s, AMs, n);
    const double omega = numerator / denominator;
    for (size_t i = 0; i < n; ++i) x[i] += alpha * Mp[i];
    for (size_t i = 0; i < n; ++i) x[i] += omega * Ms[i];
    memcpy(r, s, sizeof(double) * n);
    for (size_t i = 0; i < n; ++i) r[i] += -omega * AMs[i];
    const double conv = sqrt(orc_dot(r, r, n) * sq_inv_ini);
    PUSH(conv);
    if (conv < tol) break;
    const double beta = orc_dot(r, r0, n) * alpha / (r_r2 * omega);
    for (size_t i = 0; i < n; ++i) p[i] *= beta;
    for (size_t i = 0; i < n; ++i) p[i] += 1.0 * r[i];
    for (size_t i = 0; i < n; ++i) p[i] += -beta * omega * AMp[i];
  }
#undef PUSH
  free(s); free(Ms); free(AMs); free(Mp); free(AMp); free(r0); free(p);
  return nh;
}

/* svd3.cpp:114-194 EigenSym3: cyclic-by-largest Jacobi rotations on the
 * Frobenius-normalised symmetric matrix sm = {m00, m11, m22, m12, m20, m01}. */
static int eigen_sym3(double u[9], double l[3], const double sm[6], int nitr) {
  u[0] = u[4] = u[8] = 1.0;
  u[1] = u[2] = u[3] = u[5] = u[6] = u[7] = 0.0;
  l[0] = l[1] = l[2] = 0.0;
  const double dnrm = sm[0] * sm[0] + sm[1] * sm[1] + sm[2] * sm[2] +
                      2 * (sm[3] * sm[3] + sm[4] * sm[4] + sm[5] * sm[5]);
  if (dnrm < 1.0e-30) return 0;
  const double scale = sqrt(dnrm);
  const double invscl = 1.0 / scale;
  double sms[6] = {sm[0] * invscl, sm[1] * invscl, sm[2] * invscl,
                   sm[3] * invscl, sm[4] * invscl, sm[5] * invscl};
  for (int itr = 0; itr < nitr; itr++) {
    const double m[6] = {sms[0], sms[1], sms[2], sms[3], sms[4], sms[5]};
    const double v[9] = {u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8]};
    const double a12 = fabs(sms[3]), a20 = fabs(sms[4]), a01 = fabs(sms[5]);
    if (a12 >= a20 && a12 >= a01) {
      const double t = 0.5 * atan2(2 * m[3], m[2] - m[1]);
      const double ct = cos(t), st = sin(t);
      sms[1] = ct * ct * m[1] + st * st * m[2] - 2 * st * ct * m[3];
      sms[2] = ct * ct * m[2] + st * st * m[1] + 2 * st * ct * m[3];
      sms[3] = 0;
      sms[4] = st * m[5] + ct * m[4];
      sms[5] = ct * m[5] - st * m[4];
      u[1] = +ct * v[1] - st * v[2];
      u[2] = +st * v[1] + ct * v[2];
      u[4] = +ct * v[4] - st * v[5];
      u[5] = +st * v[4] + ct * v[5];
      u[7] = +ct * v[7] - st * v[8];
      u[8] = +st * v[7] + ct * v[8];
    } else if (a20 >= a01 && a20 >= a12) {
      const double t = 0.5 * atan2(2 * m[4], m[2] - m[0]);
      const double ct = cos(t), st = sin(t);
      sms[0] = ct * ct * m[0] + st * st * m[2] - 2 * st * ct * m[4];
      sms[2] = ct * ct * m[2] + st * st * m[0] + 2 * st * ct * m[4];
      sms[3] = st * m[5] + ct * m[3];
      sms[4] = 0;
      sms[5] = ct * m[5] - st * m[3];
      u[0] = +ct * v[0] - st * v[2];
      u[2] = +st * v[0] + ct * v[2];
      u[3] = +ct * v[3] - st * v[5];
      u[5] = +st * v[3] + ct * v[5];
      u[6] = +ct * v[6] - st * v[8];
      u[8] = +st * v[6] + ct * v[8];
    } else {
      const double t = 0.5 * atan2(2 * m[5], m[1] - m[0]);
      const double ct = cos(t), st = sin(t);
      sms[0] = ct * ct * m[0] + st * st * m[1] - 2 * st * ct * m[5];
      sms[1] = ct * ct * m[1] + st * st * m[0] + 2 * st * ct * m[5];
      sms[3] = st * m[4] + ct * m[3];
      sms[4] = ct * m[4] - st * m[3];
      sms[5] = 0;
      u[0] = +ct * v[0] - st * v[1];
      u[1] = +st * v[0] + ct * v[1];
      u[3] = +ct * v[3] - st * v[4];
      u[4] = +st * v[3] + ct * v[4];
      u[6] = +ct * v[6] - st * v[7];
      u[7] = +st * v[6] + ct * v[7];
    }
  }
  l[0] = scale * sms[0];
  l[1] = scale * sms[1];
  l[2] = scale * sms[2];
  return 1;
}

/* svd3.cpp:60-108 SortEigen3: descending eigenvalues, columns of V with them */
static void sort_eigen3(double G[3], double V[9]) {
  double t;
  if (G[1] > G[0]) {
    t = G[0]; G[0] = G[1]; G[1] = t;
    t = V[0]; V[0] = V[1]; V[1] = t;
    t = V[3]; V[3] = V[4]; V[4] = t;
    t = V[6]; V[6] = V[7]; V[7] = t;
  }
  if (G[2] > G[1]) {
    t = G[1]; G[1] = G[2]; G[2] = t;
    t = V[1]; V[1] = V[2]; V[2] = t;
    t = V[4]; V[4] = V[5]; V[5] = t;
    t = V[7]; V[7] = V[8]; V[8] = t;
  }
  if (G[1] > G[0]) {
    t = G[0]; G[0] = G[1]; G[1] = t;
    t = V[0]; V[0] = V[1]; V[1] = t;
    t = V[3]; V[3] = V[4]; V[4] = t;
    t = V[6]; V[6] = V[7]; V[7] = t;
  }
}

static void normalize3(double v[3]) {
  const double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= len;
  v[1] /= len;
  v[2] /= len;
}

/* svd3.cpp:196-289 Svd3 (m = U G V^T; only U and V are used here) and
 * svd3.cpp:291-305 GetRotPolarDecomp: R = U V^T. */
void orc_rot_polar_decomp(double R[9], const double m[9], int nitr) {
  double U[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, V[9], lv[3];
  const double mtm[6] = {m[0] * m[0] + m[3] * m[3] + m[6] * m[6],
                         m[1] * m[1] + m[4] * m[4] + m[7] * m[7],
                         m[2] * m[2] + m[5] * m[5] + m[8] * m[8],
                         m[1] * m[2] + m[4] * m[5] + m[7] * m[8],
                         m[2] * m[0] + m[5] * m[3] + m[8] * m[6],
                         m[0] * m[1] + m[3] * m[4] + m[6] * m[7]};
  eigen_sym3(V, lv, mtm, nitr);
  sort_eigen3(lv, V);
  double u0[3] = {m[0] * V[0] + m[1] * V[3] + m[2] * V[6], m[3] * V[0] + m[4] * V[3] + m[5] * V[6],
                  m[6] * V[0] + m[7] * V[3] + m[8] * V[6]};
  double u1[3] = {m[0] * V[1] + m[1] * V[4] + m[2] * V[7], m[3] * V[1] + m[4] * V[4] + m[5] * V[7],
                  m[6] * V[1] + m[7] * V[4] + m[8] * V[7]};
  double u2[3] = {m[0] * V[2] + m[1] * V[5] + m[2] * V[8], m[3] * V[2] + m[4] * V[5] + m[5] * V[8],
                  m[6] * V[2] + m[7] * V[5] + m[8] * V[8]};
  const double det = +V[0] * V[4] * V[8] + V[3] * V[7] * V[2] + V[6] * V[1] * V[5] -
                     V[0] * V[7] * V[5] - V[6] * V[4] * V[2] - V[3] * V[1] * V[8];
  if (det < 0) {  /* right-handed V (the U entries flipped here are overwritten below) */
    V[2] *= -1;
    V[5] *= -1;
    V[8] *= -1;
  }
  const double sql0 = u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2];
  if (sql0 > 1.0e-20) {
    normalize3(u0);
    const double sql1 = u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2];
    if (sql1 < 1.0e-20) {
      u1[0] = 1.0 - fabs(u0[0]);
      u1[1] = 1.0 - fabs(u0[1]);
      u1[2] = 1.0 - fabs(u0[2]);
    } else {
      normalize3(u1);
    }
    const double d01 = u0[0] * u1[0] + u0[1] * u1[1] + u0[2] * u1[2];
    u1[0] -= d01 * u0[0];
    u1[1] -= d01 * u0[1];
    u1[2] -= d01 * u0[2];
    normalize3(u1);
    double s2[3];
    s2[0] = u0[1] * u1[2] - u1[1] * u0[2];
    s2[1] = u0[2] * u1[0] - u1[2] * u0[0];
    s2[2] = u0[0] * u1[1] - u1[0] * u0[1];
    u2[0] = s2[0];
    u2[1] = s2[1];
    u2[2] = s2[2];
  } else {
    u0[0] = 1; u0[1] = 0; u0[2] = 0;
    u1[0] = 0; u1[1] = 1; u1[2] = 0;
    u2[0] = 0; u2[1] = 0; u2[2] = 1;
  }
  U[0] = u0[0]; U[1] = u1[0]; U[2] = u2[0];
  U[3] = u0[1]; U[4] = u1[1]; U[5] = u2[1];
  U[6] = u0[2]; U[7] = u1[2]; U[8] = u2[2];
  /* svd3.cpp:46-58 MatMatT3: R = U V^T */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      R[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] +
                     U[i * 3 + 2] * V[j * 3 + 2];
}

/* examples_oldgl_glfw/603_Fem3DStiffwarp/main.cpp:80-106 RotationAtMeshPoints:
 * per node, A = sum over the one-ring of (p_j - p_i) (P_j - P_i)^T with
 * p = P + u, then R_i = GetRotPolarDecomp(A, 100). */
void orc_rotation_at_points(double *R, const double *xyz, const double *disp,
                            const unsigned *psup_ind, const unsigned *psup, size_t np) {
  for (size_t ip = 0; ip < np; ++ip) {
    const double Pi[3] = {xyz[ip * 3 + 0], xyz[ip * 3 + 1], xyz[ip * 3 + 2]};
    const double pi[3] = {xyz[ip * 3 + 0] + disp[ip * 3 + 0], xyz[ip * 3 + 1] + disp[ip * 3 + 1],
                          xyz[ip * 3 + 2] + disp[ip * 3 + 2]};
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (unsigned jj = psup_ind[ip]; jj < psup_ind[ip + 1]; ++jj) {
      const unsigned jp = psup[jj];
      const double Pj[3] = {xyz[jp * 3 + 0], xyz[jp * 3 + 1], xyz[jp * 3 + 2]};
      const double pj[3] = {xyz[jp * 3 + 0] + disp[jp * 3 + 0], xyz[jp * 3 + 1] + disp[jp * 3 + 1],
                            xyz[jp * 3 + 2] + disp[jp * 3 + 2]};
      const double a[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      const double b[3] = {Pj[0] - Pi[0], Pj[1] - Pi[1], Pj[2] - Pi[2]};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i * 3 + j] += a[i] * b[j];
    }
    orc_rot_polar_decomp(R + ip * 9, A, 100);
  }
}

/* fem_solidlinear.h:445-530 MergeLinSys_SolidStiffwarp_BEuler_MeshTet3D:
 * emat_ij = R_i K_e,ij R_i^T (femutil.cpp:107-116 MatMatTrans3, then
 * :96-105 MatMat3), eres_i = vol rho g / 4 - sum_j emat_ij (P_j + u_j +
 * dt v_j - R_i P_j) (three subtractions per pair, femutil.cpp:88-94
 * MatVec3), lumped rho vol / 4 / dt^2 on the diagonal, vec_b += eres / dt. */
int orc_assemble_tet_stiffwarp(unsigned nblk, const unsigned *colind, const unsigned *rowptr,
                               double *vcrs, double *vdia, double *vec_b, const double *myu_e,
                               const double *lambda_e, double rho, const double *g, double dt,
                               const double *xyz, const unsigned *tet, size_t ntet,
                               const double *disp, const double *velo, const double *R) {
  unsigned *buf = (unsigned *)malloc(sizeof(unsigned) * (nblk + 1));
  for (unsigned i = 0; i < nblk; ++i) buf[i] = UINT_MAX;
  int status = 0;
  for (size_t iel = 0; iel < ntet; ++iel) {
    const unsigned *ip = tet + iel * 4;
    double P[4][3], u0[4][3], v0[4][3];
    fetch43(P, ip, xyz);
    const double vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
    double d[4][3], emat0[4][4][3][3], emat[4][4][3][3], eres[4][3];
    orc_tet_dldx(d, P[0], P[1], P[2], P[3]);
    set_emat_ls3(emat0, lambda_e[iel], myu_e[iel], (const double(*)[3])d, vol);
    for (int ino = 0; ino < 4; ++ino) {
      const double *Mi = R + (size_t)ip[ino] * 9;
      for (int jno = 0; jno < 4; ++jno) {
        const double *K = &emat0[ino][jno][0][0];
        double mtmp[9];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            mtmp[i * 3 + j] = K[i * 3 + 0] * Mi[j * 3 + 0] + K[i * 3 + 1] * Mi[j * 3 + 1] +
                              K[i * 3 + 2] * Mi[j * 3 + 2];
        double *E = &emat[ino][jno][0][0];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            E[i * 3 + j] = Mi[i * 3 + 0] * mtmp[0 * 3 + j] + Mi[i * 3 + 1] * mtmp[1 * 3 + j] +
                           Mi[i * 3 + 2] * mtmp[2 * 3 + j];
      }
    }
    for (int k = 0; k < 4; ++k) {
      eres[k][0] = vol * rho * g[0] * 0.25;
      eres[k][1] = vol * rho * g[1] * 0.25;
      eres[k][2] = vol * rho * g[2] * 0.25;
    }
    fetch43(u0, ip, disp);
    fetch43(v0, ip, velo);
    for (int ino = 0; ino < 4; ++ino) {
      const double *Mi = R + (size_t)ip[ino] * 9;
      for (int idim = 0; idim < 3; ++idim)
        for (int jno = 0; jno < 4; ++jno) {
          double Pj1[3];
          Pj1[0] = Mi[0] * P[jno][0] + Mi[1] * P[jno][1] + Mi[2] * P[jno][2];
          Pj1[1] = Mi[3] * P[jno][0] + Mi[4] * P[jno][1] + Mi[5] * P[jno][2];
          Pj1[2] = Mi[6] * P[jno][0] + Mi[7] * P[jno][1] + Mi[8] * P[jno][2];
          const double uj1[3] = {P[jno][0] + u0[jno][0] + dt * v0[jno][0] - Pj1[0],
                                 P[jno][1] + u0[jno][1] + dt * v0[jno][1] - Pj1[1],
                                 P[jno][2] + u0[jno][2] + dt * v0[jno][2] - Pj1[2]};
          eres[ino][idim] -= emat[ino][jno][idim][0] * uj1[0];
          eres[ino][idim] -= emat[ino][jno][idim][1] * uj1[1];
          eres[ino][idim] -= emat[ino][jno][idim][2] * uj1[2];
        }
    }
    for (int ino = 0; ino < 4; ++ino) {
      emat[ino][ino][0][0] += rho * vol * 0.25 / (dt * dt);
      emat[ino][ino][1][1] += rho * vol * 0.25 / (dt * dt);
      emat[ino][ino][2][2] += rho * vol * 0.25 / (dt * dt);
    }
    for (int k = 0; k < 4; ++k) {
      vec_b[ip[k] * 3 + 0] += eres[k][0] / dt;
      vec_b[ip[k] * 3 + 1] += eres[k][1] / dt;
      vec_b[ip[k] * 3 + 2] += eres[k][2] / dt;
    }
    status |= merge_tet(ip, (const double(*)[4][3][3])emat, colind, rowptr, vcrs, vdia, buf);
  }
  free(buf);
  return status;
}
