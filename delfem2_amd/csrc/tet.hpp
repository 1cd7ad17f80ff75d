// tet.hpp -- linear tetrahedron geometry and element stiffness entries,
// shared by the gather assembly and the adjoint contraction.
#pragma once

#include "common.hpp"

namespace dfm2h {

// ---------------------------------------------------------------------------
// tet geometry (femutil.cpp:202-258) and element-matrix entries
// (fem_solidlinear.cpp:55-79), same association as the reference.

__device__ inline double tet_volume(const double *v1, const double *v2, const double *v3,
                                    const double *v4) {
  const double one_sixth = 1.0 / 6.0;
  return ((v2[0] - v1[0]) * ((v3[1] - v1[1]) * (v4[2] - v1[2]) - (v4[1] - v1[1]) * (v3[2] - v1[2]))
          - (v2[1] - v1[1]) * ((v3[0] - v1[0]) * (v4[2] - v1[2]) - (v4[0] - v1[0]) * (v3[2] - v1[2]))
          + (v2[2] - v1[2]) * ((v3[0] - v1[0]) * (v4[1] - v1[1]) - (v4[0] - v1[0]) * (v3[1] - v1[1]))) *
         one_sixth;
}

// gradient of barycentric coordinate n: normal of the opposite face (nodes
// n+1, n+2, n+3 cyclic) over 6*vol, signs (s,-s,s) with s = -,+,-,+ over n.
__device__ inline void tet_dldx_row(double out[3], const double *qa, const double *qb,
                                    const double *qc, double s) {
  out[0] = s * ((qb[1] - qa[1]) * (qc[2] - qa[2]) - (qc[1] - qa[1]) * (qb[2] - qa[2]));
  out[1] = -s * ((qb[0] - qa[0]) * (qc[2] - qa[2]) - (qc[0] - qa[0]) * (qb[2] - qa[2]));
  out[2] = s * ((qb[0] - qa[0]) * (qc[1] - qa[1]) - (qc[0] - qa[0]) * (qb[1] - qa[1]));
}

struct TetGeo {
  double vol;
  double d[4][3];
};

__device__ inline void tet_geometry(TetGeo &g, const double P[4][3]) {
  g.vol = tet_volume(P[0], P[1], P[2], P[3]);
  const double dtmp1 = 1.0 / (g.vol * 6.0);
  tet_dldx_row(g.d[0], P[1], P[2], P[3], -dtmp1);
  tet_dldx_row(g.d[1], P[2], P[3], P[0], dtmp1);
  tet_dldx_row(g.d[2], P[3], P[0], P[1], -dtmp1);
  tet_dldx_row(g.d[3], P[0], P[1], P[2], dtmp1);
}

// emat[li][lj][a][b] built exactly as AddEmat_LinearSolid3 does on a zeroed
// emat: 0 + w*(lambda*di_a*dj_b + myu*dj_a*di_b), plus w*myu*(di.dj) when a==b
// (SET: assigned as SetEmat_LinearSolid3 does, fem_solidlinear.h:129-153,
// which keeps the sign of an exact zero).
template <int a, int b, bool SET = false>
__device__ inline double emat_entry(const double *di, const double *dj, double w, double myu,
                                    double lam) {
  double v;
  if (SET) {
    v = w * (lam * di[a] * dj[b] + myu * dj[a] * di[b]);
  } else {
    v = 0.0;
    v += w * (lam * di[a] * dj[b] + myu * dj[a] * di[b]);
  }
  if (a == b) {
    const double dt = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
    v += w * myu * dt;
  }
  return v;
}

template <bool SET = false>
__device__ inline void emat_block(double blk[9], const double *di, const double *dj, double w,
                                  double myu, double lam) {
  blk[0] = emat_entry<0, 0, SET>(di, dj, w, myu, lam);
  blk[1] = emat_entry<0, 1, SET>(di, dj, w, myu, lam);
  blk[2] = emat_entry<0, 2, SET>(di, dj, w, myu, lam);
  blk[3] = emat_entry<1, 0, SET>(di, dj, w, myu, lam);
  blk[4] = emat_entry<1, 1, SET>(di, dj, w, myu, lam);
  blk[5] = emat_entry<1, 2, SET>(di, dj, w, myu, lam);
  blk[6] = emat_entry<2, 0, SET>(di, dj, w, myu, lam);
  blk[7] = emat_entry<2, 1, SET>(di, dj, w, myu, lam);
  blk[8] = emat_entry<2, 2, SET>(di, dj, w, myu, lam);
}

__device__ inline int local_index(uint4 n, unsigned v) {
  return n.x == v ? 0 : n.y == v ? 1 : n.z == v ? 2 : n.w == v ? 3 : -1;
}

// row k (0..3, run-time) of the gradient table, selected without indexing
// private memory dynamically
__device__ inline void pick_row(double out[3], const TetGeo &g, int k) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
    out[a] = k == 0 ? g.d[0][a] : k == 1 ? g.d[1][a] : k == 2 ? g.d[2][a] : g.d[3][a];
}

__device__ inline void load_tet(double P[4][3], const double *__restrict__ xyz, uint4 n) {
  const unsigned id[4] = {n.x, n.y, n.z, n.w};
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) P[k][a] = xyz[(size_t)id[k] * 3 + a];
}

}  // namespace dfm2h
