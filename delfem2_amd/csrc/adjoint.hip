// adjoint.hip -- per-element adjoint contraction for material sensitivities.
// No reference counterpart (the reference has no adjoint).
//
// For K(E) = sum_e E_e K_e^1 (K_e^1: element stiffness at unit modulus, i.e.
// at the Lame pair (myu1_e, lambda1_e)) and an objective J(u) with K u = f,
// the adjoint lam solves K lam = dJ/du (fixed DOFs zero) and
//   dJ/dE_e = -lam_e^T K_e^1 u_e .
// One thread per element; the 16 blocks are contracted in (i, j, a, b) order
// as sum += (lam_ia * K_ijab) * u_jb, the order the CPU oracle uses, so the
// result is bit-identical to it for the same u and lam.
#include "handles.hpp"
#include "tet.hpp"

namespace dfm2h {

__global__ void __launch_bounds__(256)
k_adjoint_grad_tet(size_t ntet, const uint4 *__restrict__ tet, const double *__restrict__ xyz,
                   const double *__restrict__ myu1, const double *__restrict__ lam1,
                   const double *__restrict__ u, const double *__restrict__ lam,
                   double *__restrict__ grad) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntet) return;
  const uint4 n = tet[e];
  const unsigned id[4] = {n.x, n.y, n.z, n.w};
  double P[4][3], U[4][3], L[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      P[k][a] = xyz[(size_t)id[k] * 3 + a];
      U[k][a] = u[(size_t)id[k] * 3 + a];
      L[k][a] = lam[(size_t)id[k] * 3 + a];
    }
  TetGeo G;
  tet_geometry(G, P);
  const double my = myu1[e], la = lam1[e];
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double blk[9];
      emat_block(blk, G.d[i], G.d[j], G.vol, my, la);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) s += L[i][a] * blk[a * 3 + b] * U[j][b];
    }
  grad[e] = -s;
}

}  // namespace dfm2h

using namespace dfm2h;

extern "C" int dfm2h_adjoint_grad_tet(dfm2h_tetmesh M, const double *xyz, const double *myu1_e,
                                      const double *lambda1_e, const double *u,
                                      const double *lam, double *grad) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(M, DFM2H_ERR_INVALID, "null mesh handle");
    DFM2H_REQUIRE(xyz && myu1_e && lambda1_e && u && lam && grad, DFM2H_ERR_INVALID,
                  "null argument");
    for (const double *p : {xyz, myu1_e, lambda1_e, u, lam, (const double *)grad})
      require_device(p, "a device array argument");
    if (!M->ntet) return;
    hipLaunchKernelGGL(k_adjoint_grad_tet, dim3(ceil_div(M->ntet, 256)), dim3(256), 0, stream(),
                       M->ntet, reinterpret_cast<const uint4 *>(M->tet.p), xyz, myu1_e,
                       lambda1_e, u, lam, grad);
    check_launch("k_adjoint_grad_tet");
  });
}
