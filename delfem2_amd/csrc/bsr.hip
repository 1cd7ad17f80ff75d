// bsr.hip -- CMatrixSparse<double> (3x3 blocks, separate diagonal) on the
// device: pattern, values, MatVec, Dirichlet rows/columns, and the gather
// form of the tet assembly.
#include "handles.hpp"
#include "tet.hpp"

#include <cstdlib>
#include <mutex>
#include <set>

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
  // time-stepping drivers: step, Newmark beta, velocity and acceleration,
  // node rotations (stiffness warping, 9 per node, row-major)
  double dt, beta_nb;
  const double *velo, *acc, *rot;
};

// Element kinds of the gather assembly:
//   static    fem_solidlinear.h:192-237 (EMat_SolidLinear_Static_Tet, additive form)
//   newmark   fem_solidlinear.h:331-379 with fem_solidlinear.cpp:437-490
//             (is_initial_iter = true): M_e + beta dt^2 K_e,
//             eres = f - K_e u - M_e a - dt K_e v - (0.5 dt^2) K_e a
//   beuler    fem_solidlinear.h:381-443: K_e + rho vol / 4 / dt^2 on the
//             diagonal entries of the diagonal blocks, eres = (f - K_e (u + dt v)) / dt
//   stiffwarp fem_solidlinear.h:445-530: blocks R_i K_e,ij R_i^T (R_i: rotation
//             of the row node), eres_i = (f - sum_j E_ij (P_j + u_j + dt v_j - R_i P_j)) / dt,
//             the same lumped mass
enum AsmKind { kAsmStatic = 0, kAsmNewmark = 1, kAsmBEuler = 2, kAsmStiffwarp = 3 };

// E = R K R^T as femutil.cpp:107-116 MatMatTrans3 (K R^T) then :96-105 MatMat3
__device__ inline void rotate_block(double E[9], const double K[9], const double R[9]) {
  double t[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      t[i * 3 + j] = K[i * 3 + 0] * R[j * 3 + 0] + K[i * 3 + 1] * R[j * 3 + 1] +
                     K[i * 3 + 2] * R[j * 3 + 2];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      E[i * 3 + j] = R[i * 3 + 0] * t[0 * 3 + j] + R[i * 3 + 1] * t[1 * 3 + j] +
                     R[i * 3 + 2] * t[2 * 3 + j];
}

__device__ inline void load_node3(double out[3], const double *v, unsigned n) {
  out[0] = v ? v[(size_t)n * 3 + 0] : 0.0;
  out[1] = v ? v[(size_t)n * 3 + 1] : 0.0;
  out[2] = v ? v[(size_t)n * 3 + 2] : 0.0;
}

// R_li += scale * (E_li,jn v_jn) for one node pair (AddEmatEvecScale3<4>'s
// inner step, femutil.h:252-268)
__device__ inline void add_block_vec(double er[3], const double e[9], const double v[3],
                                     double scale) {
  er[0] += scale * (e[0] * v[0] + e[1] * v[1] + e[2] * v[2]);
  er[1] += scale * (e[3] * v[0] + e[4] * v[1] + e[5] * v[2]);
  er[2] += scale * (e[6] * v[0] + e[7] * v[1] + e[8] * v[2]);
}

// consistent mass block (li, lj) of SetEmatConsistentMassTet (femutil.cpp:260-280)
__device__ inline void mass_block(double m[9], double dm, bool same) {
  const double d = same ? dm + dm : dm;
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = (k % 4 == 0) ? d : 0.0;
}

// One element's contribution to the destination block (li, lj) of row node
// i, per kind (the reference's element matrix entry, same expressions).
template <int KIND>
__device__ inline void element_block(const AsmArgs &A, unsigned i, double vol, const double di[3],
                                     const double dj[3], bool same, double myu, double lam,
                                     double blk[9]) {
  const double w = vol;
  if (KIND == kAsmStatic) {
    emat_block(blk, di, dj, w, myu, lam);
  } else if (KIND == kAsmNewmark) {  // (&emat)[k] = eM[k] + (beta dt dt) eK[k]
    double kb[9], mb[9];
    emat_block<true>(kb, di, dj, w, myu, lam);
    mass_block(mb, A.rho * vol * 0.05, same);
    const double s = A.beta_nb * A.dt * A.dt;
#pragma unroll
    for (int k = 0; k < 9; ++k) blk[k] = mb[k] + s * kb[k];
  } else if (KIND == kAsmStiffwarp) {  // R_i K_e R_i^T, lumped mass on the diagonal block
    double kb[9], Ri[9];
    emat_block<true>(kb, di, dj, w, myu, lam);
#pragma unroll
    for (int k = 0; k < 9; ++k) Ri[k] = A.rot[(size_t)i * 9 + k];
    rotate_block(blk, kb, Ri);
    if (same) {
      blk[0] += A.rho * vol * 0.25 / (A.dt * A.dt);
      blk[4] += A.rho * vol * 0.25 / (A.dt * A.dt);
      blk[8] += A.rho * vol * 0.25 / (A.dt * A.dt);
    }
  } else {  // backward Euler: K_e, lumped mass on the diagonal block
    emat_block<true>(blk, di, dj, w, myu, lam);
    if (same) {
      blk[0] += A.rho * vol * 0.25 / (A.dt * A.dt);
      blk[4] += A.rho * vol * 0.25 / (A.dt * A.dt);
      blk[8] += A.rho * vol * 0.25 / (A.dt * A.dt);
    }
  }
}

// One element's contribution to the right-hand side of row node i (local
// index li, gradient row di): the addend the reference's per-element eres
// puts into vec_b at node i, per kind, written to out (the caller adds it).
// P: the element's node coordinates (read by the stiffness-warping kind only).
template <int KIND>
__device__ inline void element_rhs_term(const AsmArgs &A, unsigned i, int li, uint4 n,
                                        const TetGeo &G, const double (&P)[4][3],
                                        const double di[3], double myu, double lam,
                                        double out[3]) {
  const double w = G.vol;
  const double vol = G.vol;
  const unsigned id[4] = {n.x, n.y, n.z, n.w};
  if (KIND == kAsmNewmark) {
    double er[3] = {vol * A.rho * A.g0 * 0.25, vol * A.rho * A.g1 * 0.25,
                    vol * A.rho * A.g2 * 0.25};
    double kb[9], mb[9], x[3];
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {  // - K_e u
      emat_block<true>(kb, di, G.d[jn], w, myu, lam);
      load_node3(x, A.disp, id[jn]);
      add_block_vec(er, kb, x, -1.);
    }
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {  // - M_e a
      mass_block(mb, A.rho * vol * 0.05, jn == li);
      load_node3(x, A.acc, id[jn]);
      add_block_vec(er, mb, x, -1.);
    }
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {  // - dt K_e v
      emat_block<true>(kb, di, G.d[jn], w, myu, lam);
      load_node3(x, A.velo, id[jn]);
      add_block_vec(er, kb, x, -A.dt);
    }
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {  // - (0.5 dt^2) K_e a
      emat_block<true>(kb, di, G.d[jn], w, myu, lam);
      load_node3(x, A.acc, id[jn]);
      add_block_vec(er, kb, x, -0.5 * A.dt * A.dt);
    }
    out[0] = er[0];
    out[1] = er[1];
    out[2] = er[2];
  } else if (KIND == kAsmBEuler) {
    double er[3] = {vol * A.rho * A.g0 * 0.25, vol * A.rho * A.g1 * 0.25,
                    vol * A.rho * A.g2 * 0.25};
    double kb[9], u[3], v[3];
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {  // - K_e (u + dt v), three subtractions per pair
      emat_block<true>(kb, di, G.d[jn], w, myu, lam);
      load_node3(u, A.disp, id[jn]);
      load_node3(v, A.velo, id[jn]);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        er[a] -= kb[3 * a + 0] * (u[0] + A.dt * v[0]);
        er[a] -= kb[3 * a + 1] * (u[1] + A.dt * v[1]);
        er[a] -= kb[3 * a + 2] * (u[2] + A.dt * v[2]);
      }
    }
    out[0] = er[0] / A.dt;
    out[1] = er[1] / A.dt;
    out[2] = er[2] / A.dt;
  } else if (KIND == kAsmStiffwarp) {
    double er[3] = {vol * A.rho * A.g0 * 0.25, vol * A.rho * A.g1 * 0.25,
                    vol * A.rho * A.g2 * 0.25};
    double Ri[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ri[k] = A.rot[(size_t)i * 9 + k];
#pragma unroll
    for (int jn = 0; jn < 4; ++jn) {
      double kb[9], E[9], u[3], v[3];
      emat_block<true>(kb, di, G.d[jn], w, myu, lam);
      rotate_block(E, kb, Ri);
      load_node3(u, A.disp, id[jn]);
      load_node3(v, A.velo, id[jn]);
      const double *Pj = P[jn];
      const double Pj1[3] = {Ri[0] * Pj[0] + Ri[1] * Pj[1] + Ri[2] * Pj[2],
                             Ri[3] * Pj[0] + Ri[4] * Pj[1] + Ri[5] * Pj[2],
                             Ri[6] * Pj[0] + Ri[7] * Pj[1] + Ri[8] * Pj[2]};
      const double uj1[3] = {Pj[0] + u[0] + A.dt * v[0] - Pj1[0],
                             Pj[1] + u[1] + A.dt * v[1] - Pj1[1],
                             Pj[2] + u[2] + A.dt * v[2] - Pj1[2]};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        er[a] -= E[3 * a + 0] * uj1[0];
        er[a] -= E[3 * a + 1] * uj1[1];
        er[a] -= E[3 * a + 2] * uj1[2];
      }
    }
    out[0] = er[0] / A.dt;
    out[1] = er[1] / A.dt;
    out[2] = er[2] / A.dt;
  } else {
    // eres = vol*rho*g*0.25 - K_e u_e (fem_solidlinear.h:214-221 and
    // AddEmatEvecScale3 femutil.h:252-268, scale -1), then vec_b += eres
    double er0 = vol * A.rho * A.g0 * 0.25;
    double er1 = vol * A.rho * A.g1 * 0.25;
    double er2 = vol * A.rho * A.g2 * 0.25;
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
    out[0] = er0;
    out[1] = er1;
    out[2] = er2;
  }
}

// every neighbour of row node i in an element must own a slot in row i
__device__ inline void check_row_slots(const AsmArgs &A, unsigned i, uint4 n) {
  const unsigned id[4] = {n.x, n.y, n.z, n.w};
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

// Gather assembly, one thread per destination block: off-diagonal slot s
// (t < ncrs) or the diagonal block and right-hand side of row t-ncrs.  Each
// thread walks the elements around its row node in ascending element order
// and adds their contribution to its block; the additions therefore happen
// in exactly the order of the reference's sequential element loop with
// Merge<4,4,3,3>, and no two threads write the same address (no atomics).
// This form computes each element's geometry where it is used; the listed
// form below is the one run, this one serves meshes of 2^28 tets or more.
template <int KIND>
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
    double di[3], dj[3];
    pick_row(di, G, li);
    pick_row(dj, G, lj);
    double blk[9];
    element_block<KIND>(A, i, G.vol, di, dj, li == lj, myu, lam, blk);
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] += blk[k];
    if (diag) {
      double tr[3];
      element_rhs_term<KIND>(A, i, li, n, G, P, di, myu, lam, tr);
      bacc[0] += tr[0];
      bacc[1] += tr[1];
      bacc[2] += tr[2];
      check_row_slots(A, i, n);
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

// Per-tet record of the listed assembly: volume, the four gradient rows and
// the element's Lame parameters (tet_geometry's values).
constexpr int kTetRec = 16;  // doubles: vol, d[4][3], myu, lambda, unused

// (256 threads; the workgroup's records go through LDS so that the global
// stores are contiguous 16-byte runs)
__global__ void __launch_bounds__(256) k_tet_records(AsmArgs A, size_t ntet, double *rec) {
  __shared__ double lr[256][kTetRec + 1];  // padded against bank conflicts
  const size_t e0 = (size_t)wg_id(true) * 256, e = e0 + threadIdx.x;
  if (e < ntet) {
    double P[4][3];
    load_tet(P, A.xyz, A.tet[e]);
    TetGeo G;
    tet_geometry(G, P);
    double *r = lr[threadIdx.x];
    r[0] = G.vol;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int a = 0; a < 3; ++a) r[1 + 3 * k + a] = G.d[k][a];
    r[13] = A.myu_e ? A.myu_e[e] : A.myu;
    r[14] = A.lambda_e ? A.lambda_e[e] : A.lambda;
    r[15] = 0.0;
  }
  __syncthreads();
  const size_t n = min((size_t)256, ntet - e0) * (kTetRec / 2);  // 16-byte pieces
  double2 *out = reinterpret_cast<double2 *>(rec + e0 * kTetRec);
  for (size_t j = threadIdx.x; j < n; j += 256) {
    const size_t rr = j / (kTetRec / 2), f = 2 * (j % (kTetRec / 2));
    out[j] = double2{lr[rr][f], lr[rr][f + 1]};
  }
}

// Element entry of destination block t: the elements k_assemble_tet_gather
// would add there, in its order, packed as p | li << 28 | lj << 30 where p is
// the element id for a diagonal block (t >= ncrs: the row node's whole
// element list) and, for an off-diagonal block, the element's position in
// its row node's list (so a row's records can be staged once).  ptr ==
// nullptr: count them into cnt[t] (and flag missing slots); else write them
// from ptr[t].
__global__ void __launch_bounds__(256)
k_asm_lists(AsmArgs A, const unsigned *__restrict__ ptr, unsigned *out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool diag = t >= A.ncrs;
  if (t >= A.ncrs + A.nblk) return;
  const unsigned i = diag ? (unsigned)(t - A.ncrs) : A.rowof[t];
  const unsigned j = diag ? i : A.rowptr[t];
  unsigned c = 0;
  const size_t w = ptr ? ptr[t] : 0;
  const unsigned q1 = A.elsup_ind[i + 1];
  for (unsigned q = A.elsup_ind[i]; q < q1; ++q) {
    const unsigned e = A.elsup[q];
    const uint4 n = A.tet[e];
    const int li = local_index(n, i);
    const int lj = diag ? li : local_index(n, j);
    if (lj < 0) continue;
    if (ptr)
      out[w + c] = (diag ? e : q - A.elsup_ind[i]) | ((unsigned)li & 3u) << 28 |
                   ((unsigned)lj & 3u) << 30;
    else if (diag)
      check_row_slots(A, i, n);
    ++c;
  }
  if (!ptr) out[t] = c;
}

// The gather assembly over the element lists, off-diagonal blocks (slots
// t < ncrs): the same additions in the same order as k_assemble_tet_gather,
// each element's geometry read from its record instead of recomputed, and
// no walk over the elements that do not touch the block.  Workgroups are
// numbered per XCD (wg_id): each XCD's L2 then serves one contiguous run of
// rows and the records their elements share.
template <int KIND>
__global__ void __launch_bounds__(256)
k_assemble_offdiag_lists(AsmArgs A, const unsigned *__restrict__ ptr,
                         const unsigned *__restrict__ ent, const double *__restrict__ rec) {
  const size_t t = (size_t)wg_id(true) * blockDim.x + threadIdx.x;
  if (t >= A.ncrs) return;
  const unsigned i = A.rowof[t];
  double *dst = A.vcrs + t * 9;
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = dst[q];
  const unsigned *row_ent = ent + ptr[A.ncrs + i];  // the row node's elements
  // elements in rounds of U: the entries, then their records, are requested
  // together, then added in order
  constexpr int U = 4;
  const unsigned q0 = ptr[t], q1 = ptr[t + 1];
  for (unsigned qb = q0; qb < q1; qb += U) {
    unsigned en[U];
#pragma unroll
    for (int u = 0; u < U; ++u) en[u] = qb + u < q1 ? ent[qb + u] : 0u;
    double vol[U], my[U], la[U], di[U][3], dj[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (qb + u < q1) {
        const unsigned li = (en[u] >> 28) & 3u, lj = en[u] >> 30;
        const unsigned e = row_ent[en[u] & 0x0FFFFFFFu] & 0x0FFFFFFFu;
        const double *r = rec + (size_t)e * kTetRec;
        vol[u] = r[0];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          di[u][a] = r[1 + 3 * li + a];
          dj[u][a] = r[1 + 3 * lj + a];
        }
        my[u] = r[13];
        la[u] = r[14];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (qb + u < q1) {
        double blk[9];
        element_block<KIND>(A, i, vol[u], di[u], dj[u], ((en[u] >> 28) & 3u) == (en[u] >> 30),
                            my[u], la[u], blk);
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] += blk[k];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) dst[q] = acc[q];
}

// Off-diagonal blocks, 32 lanes per row node: the records of the node's
// elements (its diagonal list, at most kRowElems) are staged in LDS with
// 16-byte loads, then lane l adds the elements of the row's slots l, l + 32,
// ... from there -- the additions of k_assemble_offdiag_lists, with a record
// read from HBM once per row instead of nine scattered loads per block.
constexpr unsigned kRowElems = 32;
constexpr unsigned kOffdiagBatch = 4;  // list entries of a block loaded together

template <int KIND>
__global__ void __launch_bounds__(64)
k_assemble_offdiag_rows(AsmArgs A, const unsigned *__restrict__ ptr,
                        const unsigned *__restrict__ ent, const double *__restrict__ rec) {
  __shared__ double recs[2][kRowElems][kTetRec];
  const unsigned lane = threadIdx.x, grp = lane / 32, l = lane & 31u;
  const unsigned i = wg_id(true) * 2 + grp;
  const bool live = i < A.nblk;
  const unsigned d0 = live ? ptr[A.ncrs + i] : 0u, d1 = live ? ptr[A.ncrs + i + 1] : 0u;
  const unsigned m = min(d1 - d0, kRowElems);  // the host guarantees d1 - d0 <= kRowElems
  // this lane's first block: its value and element range are requested
  // together with the records
  const unsigned s0 = live ? A.colind[i] : 0u, s1 = live ? A.colind[i + 1] : 0u;
  double acc0[9];
  unsigned qa0 = 0, qb0 = 0;
  if (s0 + l < s1) {
    const double *d = A.vcrs + (size_t)(s0 + l) * 9;
#pragma unroll
    for (int q = 0; q < 9; ++q) acc0[q] = d[q];
    qa0 = ptr[s0 + l];
    qb0 = ptr[s0 + l + 1];
  }
  // the node's records into LDS, 16 bytes per lane and step: every step's
  // list entry is requested first, then every step's record part, so the
  // steps' loads are in flight together
  constexpr unsigned kParts = kTetRec / 2, kSteps = kRowElems * kParts / 32;
  unsigned es[kSteps];
#pragma unroll
  for (unsigned st = 0; st < kSteps; ++st) {
    const unsigned k = l + 32 * st;
    es[st] = k < m * kParts ? ent[d0 + k / kParts] & 0x0FFFFFFFu : 0u;
  }
  double2 vs[kSteps];
#pragma unroll
  for (unsigned st = 0; st < kSteps; ++st) {
    const unsigned k = l + 32 * st;
    if (k < m * kParts)
      vs[st] = *reinterpret_cast<const double2 *>(rec + (size_t)es[st] * kTetRec + 2 * (k % kParts));
  }
#pragma unroll
  for (unsigned st = 0; st < kSteps; ++st) {
    const unsigned k = l + 32 * st;
    if (k < m * kParts) {
      recs[grp][k / kParts][2 * (k % kParts)] = vs[st].x;
      recs[grp][k / kParts][2 * (k % kParts) + 1] = vs[st].y;
    }
  }
  __syncthreads();
  for (unsigned sl = s0 + l; sl < s1; sl += 32) {
    double *dst = A.vcrs + (size_t)sl * 9;
    const bool first = sl == s0 + l;
    double acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = first ? acc0[q] : dst[q];
    const unsigned q1 = first ? qb0 : ptr[sl + 1];
    // the block's list entries four at a time, their loads in flight
    // together; the elements are still added in list order
    for (unsigned q = first ? qa0 : ptr[sl]; q < q1; q += kOffdiagBatch) {
      unsigned ens[kOffdiagBatch];
#pragma unroll
      for (unsigned u = 0; u < kOffdiagBatch; ++u) ens[u] = q + u < q1 ? ent[q + u] : 0u;
#pragma unroll
      for (unsigned u = 0; u < kOffdiagBatch; ++u) {
        if (q + u >= q1) break;
        const unsigned en = ens[u];
        const unsigned li = (en >> 28) & 3u, lj = en >> 30;
        const double *r = recs[grp][en & 0x0FFFFFFFu];
        const double di[3] = {r[1 + 3 * li], r[2 + 3 * li], r[3 + 3 * li]};
        const double dj[3] = {r[1 + 3 * lj], r[2 + 3 * lj], r[3 + 3 * lj]};
        double blk[9];
        element_block<KIND>(A, i, r[0], di, dj, li == lj, r[13], r[14], blk);
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] += blk[k];
      }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = acc[q];
  }
}

// Diagonal blocks and right-hand side over the element lists: 32 lanes per
// row node, each computing one element's diagonal block and right-hand-side
// addend (12 values, into LDS); then lane q < 9 adds value q of every element
// to the diagonal entry q, and lanes 9..11 the addends to vec_b, in element
// order -- the additions and order of k_assemble_tet_gather's diagonal
// thread, with the elements' work spread over the lanes.  Rows with more
// than 32 elements take several rounds.
constexpr int kDiagLanes = 32;

template <int KIND>
__global__ void __launch_bounds__(64, 2)
k_assemble_diag_lists(AsmArgs A, const unsigned *__restrict__ ptr,
                      const unsigned *__restrict__ ent, const double *__restrict__ rec) {
  __shared__ double term[64][13];  // 12 values per lane, padded against bank conflicts
  const unsigned lane = threadIdx.x, g0 = lane & ~(kDiagLanes - 1u), l = lane & (kDiagLanes - 1u);
  const unsigned i = wg_id(true) * (64 / kDiagLanes) + lane / kDiagLanes;
  const bool live = i < A.nblk;
  const size_t t = A.ncrs + (live ? i : 0u);
  const unsigned q0 = live ? ptr[t] : 0u, q1 = live ? ptr[t + 1] : 0u;
  // the running sum of lane l < 12: diagonal entry l, or vec_b component l - 9
  double sum = 0.0;
  if (live && l < 9) sum = A.vdia[(size_t)i * 9 + l];
  if (live && l >= 9 && l < 12) sum = A.vec_b[(size_t)i * 3 + (l - 9)];
  unsigned rounds = q1 - q0;  // the wavefront's longest list, in rounds of 32
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    rounds = max(rounds, (unsigned)__shfl_xor((int)rounds, off, 64));
  rounds = (rounds + kDiagLanes - 1) / kDiagLanes;
  for (unsigned c = 0; c < rounds; ++c) {
    const unsigned qb = q0 + c * kDiagLanes, q = qb + l;
    if (q < q1) {
      const unsigned en = ent[q];
      const unsigned e = en & 0x0FFFFFFFu;
      const int li = (int)((en >> 28) & 3u);
      const double *r = rec + (size_t)e * kTetRec;
      TetGeo G;
      G.vol = r[0];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) G.d[k][a] = r[1 + 3 * k + a];
      const double myu = r[13], lam = r[14];
      double di[3];
      pick_row(di, G, li);
      double blk[9], tr[3];
      element_block<KIND>(A, i, G.vol, di, di, true, myu, lam, blk);
      const uint4 n = A.tet[e];
      double P[4][3];
      if (KIND == kAsmStiffwarp) load_tet(P, A.xyz, n);
      element_rhs_term<KIND>(A, i, li, n, G, P, di, myu, lam, tr);
#pragma unroll
      for (int k = 0; k < 9; ++k) term[lane][k] = blk[k];
      term[lane][9] = tr[0];
      term[lane][10] = tr[1];
      term[lane][11] = tr[2];
    }
    __syncthreads();
    if (l < 12 && qb < q1) {
      const unsigned m = min(q1 - qb, (unsigned)kDiagLanes);
      for (unsigned k = 0; k < m; ++k) sum += term[g0 + k][l];
    }
    __syncthreads();
  }
  if (live && l < 9) A.vdia[(size_t)i * 9 + l] = sum;
  if (live && l >= 9 && l < 12) A.vec_b[(size_t)i * 3 + (l - 9)] = sum;
}

// The Newmark-beta diagonal blocks and right-hand side with four lanes per
// element, one per element node jn (eight elements per 32-lane row group and
// round): lane (element, jn) forms K_e(li, jn) and M_e(li, jn) and the four
// products the right-hand side needs at node jn -- K u, M a, K v, K a, each
// row in add_block_vec's association -- and the lane with jn = li also the
// element's diagonal block.  Twelve lanes then apply them in the order of
// element_rhs_term: per element eres = body force, then -1 (K u) over jn,
// -1 (M a) over jn, -dt (K v) over jn, -(0.5 dt^2) (K a) over jn; and the
// diagonal block and vec_b receive the elements in ascending order.  Same
// values as k_assemble_diag_lists<kAsmNewmark>, with a quarter of its
// per-lane work and registers.
__global__ void __launch_bounds__(64, 5)
k_assemble_diag_newmark(AsmArgs A, const unsigned *__restrict__ ptr,
                        const unsigned *__restrict__ ent, const double *__restrict__ rec) {
  constexpr unsigned EPR = kDiagLanes / 4;  // elements per round
  __shared__ double prod[64][13];           // 4 products x 3 rows per lane, padded
  __shared__ double dblk[2 * EPR][10];      // per element of the round: its block, its volume
  __shared__ double eres[2 * EPR][3];       // per element of the round: its vec_b addend
  const unsigned lane = threadIdx.x, grp = lane / kDiagLanes, g0 = grp * kDiagLanes;
  const unsigned l = lane & (kDiagLanes - 1u), u = l / 4, jn = l % 4;
  const unsigned i = wg_id(true) * (64 / kDiagLanes) + grp;
  const bool live = i < A.nblk;
  const size_t t = A.ncrs + (live ? i : 0u);
  const unsigned q0 = live ? ptr[t] : 0u, q1 = live ? ptr[t + 1] : 0u;
  double sum = 0.0;  // lane l < 9: diagonal entry l; 9..11: vec_b component l - 9
  if (live && l < 9) sum = A.vdia[(size_t)i * 9 + l];
  if (live && l >= 9 && l < 12) sum = A.vec_b[(size_t)i * 3 + (l - 9)];
  unsigned rounds = q1 - q0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    rounds = max(rounds, (unsigned)__shfl_xor((int)rounds, off, 64));
  rounds = (rounds + EPR - 1) / EPR;
  // the row's first 32 list entries, one per lane of the row group: the
  // rounds take them by shuffle instead of a dependent load each
  const unsigned ent_l = q0 + l < q1 ? ent[q0 + l] : 0u;
  // this lane's list entry of round c (~0u past the list; a wavefront-uniform
  // call) and its element's node jn, requested a round ahead
  auto entry = [&](unsigned c) -> unsigned {
    const unsigned src = c * EPR + u, q = q0 + src;
    const unsigned en_s = (unsigned)__shfl((int)ent_l, (int)(g0 + (src & (kDiagLanes - 1u))), 64);
    return q < q1 ? (src < kDiagLanes ? en_s : ent[q]) : ~0u;
  };
  const unsigned *const tet_nodes = reinterpret_cast<const unsigned *>(A.tet);
  auto node_of = [&](unsigned en) -> unsigned {
    return en == ~0u ? 0u : tet_nodes[(size_t)(en & 0x0FFFFFFFu) * 4 + jn];
  };
  unsigned en_next = rounds > 0 ? entry(0) : ~0u;
  unsigned node_next = node_of(en_next);
  for (unsigned c = 0; c < rounds; ++c) {
    const unsigned qb = q0 + c * EPR;
    const unsigned en = en_next, node = node_next;
    if (c + 1 < rounds) {
      en_next = entry(c + 1);
      node_next = node_of(en_next);
    }
    if (en != ~0u) {
      const unsigned e = en & 0x0FFFFFFFu;
      const unsigned li = (en >> 28) & 3u;
      const double *r = rec + (size_t)e * kTetRec;
      const double vol = r[0], myu = r[13], lam = r[14];
      const double di[3] = {r[1 + 3 * li], r[2 + 3 * li], r[3 + 3 * li]};
      const double dj[3] = {r[1 + 3 * jn], r[2 + 3 * jn], r[3 + 3 * jn]};
      double kb[9], mb[9], xu[3], xv[3], xa[3];
      emat_block<true>(kb, di, dj, vol, myu, lam);
      mass_block(mb, A.rho * vol * 0.05, jn == li);
      load_node3(xu, A.disp, node);
      load_node3(xv, A.velo, node);
      load_node3(xa, A.acc, node);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        prod[lane][0 + a] = kb[3 * a + 0] * xu[0] + kb[3 * a + 1] * xu[1] + kb[3 * a + 2] * xu[2];
        prod[lane][3 + a] = mb[3 * a + 0] * xa[0] + mb[3 * a + 1] * xa[1] + mb[3 * a + 2] * xa[2];
        prod[lane][6 + a] = kb[3 * a + 0] * xv[0] + kb[3 * a + 1] * xv[1] + kb[3 * a + 2] * xv[2];
        prod[lane][9 + a] = kb[3 * a + 0] * xa[0] + kb[3 * a + 1] * xa[1] + kb[3 * a + 2] * xa[2];
      }
      if (jn == li) {  // (&emat)[k] = eM[k] + (beta dt dt) eK[k] of the block (li, li)
        const double sc = A.beta_nb * A.dt * A.dt;
#pragma unroll
        for (int k = 0; k < 9; ++k) dblk[grp * EPR + u][k] = mb[k] + sc * kb[k];
        dblk[grp * EPR + u][9] = vol;
      }
    }
    __syncthreads();
    if (l < 3 * EPR && qb + l / 3 < q1) {  // element k = l / 3, row a = l % 3: its eres_a
      const unsigned k = l / 3, a = l % 3, b = g0 + 4 * k;
      const double g = a == 0 ? A.g0 : a == 1 ? A.g1 : A.g2;
      double er = dblk[grp * EPR + k][9] * A.rho * g * 0.25;
      for (int j = 0; j < 4; ++j) er += -1. * prod[b + j][0 + a];
      for (int j = 0; j < 4; ++j) er += -1. * prod[b + j][3 + a];
      for (int j = 0; j < 4; ++j) er += -A.dt * prod[b + j][6 + a];
      for (int j = 0; j < 4; ++j) er += -0.5 * A.dt * A.dt * prod[b + j][9 + a];
      eres[grp * EPR + k][a] = er;
    }
    __syncthreads();
    if (l < 12 && qb < q1) {  // the elements in ascending order
      const unsigned m = min(q1 - qb, EPR);
      for (unsigned k = 0; k < m; ++k)
        sum += l < 9 ? dblk[grp * EPR + k][l] : eres[grp * EPR + k][l - 9];
    }
    __syncthreads();
  }
  if (live && l < 9) A.vdia[(size_t)i * 9 + l] = sum;
  if (live && l >= 9 && l < 12) A.vec_b[(size_t)i * 3 + (l - 9)] = sum;
}

// The stiffness-warped diagonal blocks and right-hand side with four lanes
// per element, one per element node jn: lane (element, jn) forms E = R_i
// K_e(li, jn) R_i^T, the node's u1 = P_j + u_j + dt v_j - R_i P_j and the
// nine products E[a][k] u1[k] the right-hand side subtracts, and the lane
// with jn = li also the element's diagonal block (that E plus the lumped
// mass).  Three lanes per element then subtract the products in the order
// of element_rhs_term (node by node, component by component) and divide by
// dt, and the diagonal block and vec_b receive the elements in ascending
// order.  Same values as k_assemble_diag_lists<kAsmStiffwarp>, with a
// quarter of its per-lane work.
__global__ void __launch_bounds__(64, 4)
k_assemble_diag_stiffwarp(AsmArgs A, const unsigned *__restrict__ ptr,
                          const unsigned *__restrict__ ent, const double *__restrict__ rec) {
  constexpr unsigned EPR = kDiagLanes / 4;  // elements per round
  __shared__ double prod[64][10];           // 9 products per lane, padded
  __shared__ double dblk[2 * EPR][10];      // per element of the round: its block, its volume
  __shared__ double eres[2 * EPR][3];       // per element of the round: its vec_b addend
  const unsigned lane = threadIdx.x, grp = lane / kDiagLanes, g0 = grp * kDiagLanes;
  const unsigned l = lane & (kDiagLanes - 1u), u = l / 4, jn = l % 4;
  const unsigned i = wg_id(true) * (64 / kDiagLanes) + grp;
  const bool live = i < A.nblk;
  const size_t t = A.ncrs + (live ? i : 0u);
  const unsigned q0 = live ? ptr[t] : 0u, q1 = live ? ptr[t + 1] : 0u;
  double sum = 0.0;  // lane l < 9: diagonal entry l; 9..11: vec_b component l - 9
  if (live && l < 9) sum = A.vdia[(size_t)i * 9 + l];
  if (live && l >= 9 && l < 12) sum = A.vec_b[(size_t)i * 3 + (l - 9)];
  double Ri[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Ri[k] = live ? A.rot[(size_t)i * 9 + k] : 0.0;
  unsigned rounds = q1 - q0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    rounds = max(rounds, (unsigned)__shfl_xor((int)rounds, off, 64));
  rounds = (rounds + EPR - 1) / EPR;
  const unsigned *const tet_nodes = reinterpret_cast<const unsigned *>(A.tet);
  for (unsigned c = 0; c < rounds; ++c) {
    const unsigned qb = q0 + c * EPR, q = qb + u;
    if (q < q1) {
      const unsigned en = ent[q];
      const unsigned e = en & 0x0FFFFFFFu;
      const unsigned li = (en >> 28) & 3u;
      const unsigned node = tet_nodes[(size_t)e * 4 + jn];
      const double *r = rec + (size_t)e * kTetRec;
      const double vol = r[0], myu = r[13], lam = r[14];
      const double di[3] = {r[1 + 3 * li], r[2 + 3 * li], r[3 + 3 * li]};
      const double dj[3] = {r[1 + 3 * jn], r[2 + 3 * jn], r[3 + 3 * jn]};
      double kb[9], E[9], xu[3], xv[3], Pj[3];
      emat_block<true>(kb, di, dj, vol, myu, lam);
      rotate_block(E, kb, Ri);
      load_node3(xu, A.disp, node);
      load_node3(xv, A.velo, node);
      load_node3(Pj, A.xyz, node);
      const double Pj1[3] = {Ri[0] * Pj[0] + Ri[1] * Pj[1] + Ri[2] * Pj[2],
                             Ri[3] * Pj[0] + Ri[4] * Pj[1] + Ri[5] * Pj[2],
                             Ri[6] * Pj[0] + Ri[7] * Pj[1] + Ri[8] * Pj[2]};
      const double uj1[3] = {Pj[0] + xu[0] + A.dt * xv[0] - Pj1[0],
                             Pj[1] + xu[1] + A.dt * xv[1] - Pj1[1],
                             Pj[2] + xu[2] + A.dt * xv[2] - Pj1[2]};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) prod[lane][3 * a + k] = E[3 * a + k] * uj1[k];
      if (jn == li) {  // element_block<kAsmStiffwarp> of the block (li, li)
        const double m = A.rho * vol * 0.25 / (A.dt * A.dt);
        E[0] += m;
        E[4] += m;
        E[8] += m;
#pragma unroll
        for (int k = 0; k < 9; ++k) dblk[grp * EPR + u][k] = E[k];
        dblk[grp * EPR + u][9] = vol;
      }
    }
    __syncthreads();
    if (l < 3 * EPR && qb + l / 3 < q1) {  // element k = l / 3, row a = l % 3: its eres_a
      const unsigned k = l / 3, a = l % 3, b = g0 + 4 * k;
      const double g = a == 0 ? A.g0 : a == 1 ? A.g1 : A.g2;
      double er = dblk[grp * EPR + k][9] * A.rho * g * 0.25;
      for (int j = 0; j < 4; ++j) {
        er -= prod[b + j][3 * a + 0];
        er -= prod[b + j][3 * a + 1];
        er -= prod[b + j][3 * a + 2];
      }
      eres[grp * EPR + k][a] = er / A.dt;
    }
    __syncthreads();
    if (l < 12 && qb < q1) {  // the elements in ascending order
      const unsigned m = min(q1 - qb, EPR);
      for (unsigned k = 0; k < m; ++k)
        sum += l < 9 ? dblk[grp * EPR + k][l] : eres[grp * EPR + k][l - 9];
    }
    __syncthreads();
  }
  if (live && l < 9) A.vdia[(size_t)i * 9 + l] = sum;
  if (live && l >= 9 && l < 12) A.vec_b[(size_t)i * 3 + (l - 9)] = sum;
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
  bool fri[3], fcj[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    fri[a] = flag[(size_t)i * 3 + a] != 0;
    fcj[a] = flag[(size_t)j * 3 + a] != 0;
  }
  if (!(fri[0] || fri[1] || fri[2] || fcj[0] || fcj[1] || fcj[2])) return;  // block untouched
  double *v = diag ? vdia + (size_t)i * 9 : vcrs + t * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const bool fr = fri[a];
      const bool fc = fcj[b];
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

// Staged Merge additions: one thread per (destination, component); run r =
// {dst, first, count} adds blocks first .. first+count-1 (call order) to
// its destination one after another -- the reference's `+=` sequence.
__global__ void k_merge_apply(size_t nrun, const unsigned *__restrict__ run, size_t ncrs,
                              const double *__restrict__ val, double *vcrs, double *vdia) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrun * 9) return;
  const size_t r = t / 9;
  const unsigned k = (unsigned)(t - r * 9);
  const size_t dst = run[3 * r], first = run[3 * r + 1], cnt = run[3 * r + 2];
  double *out = dst < ncrs ? vcrs + 9 * dst + k : vdia + 9 * (dst - ncrs) + k;
  double acc = *out;
  for (size_t c = 0; c < cnt; ++c) acc += val[9 * (first + c) + k];
  *out = acc;
}

namespace {
std::mutex g_live_mu;
std::set<const dfm2h_bsr_s *> g_live;  // matrix handles not yet destroyed
}  // namespace

bool bsr_alive(const dfm2h_bsr_s *A) {
  std::lock_guard<std::mutex> lk(g_live_mu);
  return g_live.count(A) != 0;
}

// Blocks staged beyond this count are applied before more are staged.
constexpr size_t kMergeChunk = size_t(1) << 21;

void flush_merge(dfm2h_bsr_s *A) {
  const size_t ne = A->mg_dst.size();
  if (!ne) return;
  // stable counting sort of the staged blocks by destination
  const size_t nd = A->pat.ncrs + A->nblk;
  std::vector<unsigned> start(nd + 1, 0);
  for (size_t e = 0; e < ne; ++e) start[A->mg_dst[e] + 1]++;
  for (size_t d = 0; d < nd; ++d) start[d + 1] += start[d];
  std::vector<unsigned> run;  // {dst, first, count} per touched destination
  for (size_t d = 0; d < nd; ++d)
    if (start[d + 1] > start[d])
      run.insert(run.end(), {(unsigned)d, start[d], start[d + 1] - start[d]});
  std::vector<double> val(9 * ne);
  for (size_t e = 0; e < ne; ++e) {
    const unsigned pos = start[A->mg_dst[e]]++;
    std::memcpy(&val[9 * (size_t)pos], &A->mg_val[9 * e], 9 * sizeof(double));
  }
  if (A->mg_val_d.n < val.size()) A->mg_val_d.alloc(val.size());
  if (A->mg_run_d.n < run.size()) A->mg_run_d.alloc(run.size());
  DFM2H_HIP(hipMemcpyAsync(A->mg_val_d.p, val.data(), val.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream()));
  DFM2H_HIP(hipMemcpyAsync(A->mg_run_d.p, run.data(), run.size() * sizeof(unsigned),
                           hipMemcpyHostToDevice, stream()));
  const size_t nrun = run.size() / 3;
  hipLaunchKernelGGL(k_merge_apply, dim3(ceil_div(nrun * 9, 256)), dim3(256), 0, stream(), nrun,
                     A->mg_run_d.p, A->pat.ncrs, A->mg_val_d.p, A->vcrs.p, A->vdia.p);
  check_launch("k_merge_apply");
  // the host arrays above are released on return: wait for the copies
  DFM2H_HIP(hipStreamSynchronize(stream()));
  A->mg_dst.clear();
  A->mg_val.clear();
  A->version++;
}

}  // namespace dfm2h

using namespace dfm2h;

static void need_pattern(dfm2h_bsr A) {
  DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
  DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_STATE, "SetPattern has not been called");
  flush_merge(A);
}

// The walking form runs for meshes whose element ids do not fit the packed
// entries (2^28 tets or more), or when the environment variable
// DFM2H_ASSEMBLY_WALK is 1 (read at every call; lets the two forms be
// compared in one process).
static bool walk_assembly(dfm2h_tetmesh M) {
  const char *v = std::getenv("DFM2H_ASSEMBLY_WALK");
  return (v && v[0] == '1') || M->ntet >= (1ull << 28);
}

// Shared host side of the gather assembly: checks, argument block, launch,
// missing-slot report.  `a` carries the kind's extra fields.
static void run_assembly(int kind, dfm2h_bsr A, double *vec_b, const double *myu_e,
                         const double *lambda_e, double myu, double lambda, double rho,
                         const double *g, const double *xyz, dfm2h_tetmesh M, AsmArgs a) {
  need_pattern(A);
  DFM2H_REQUIRE(M, DFM2H_ERR_INVALID, "null mesh handle");
  DFM2H_REQUIRE(vec_b && g && xyz, DFM2H_ERR_INVALID, "null argument");
  for (const double *p : {(const double *)vec_b, myu_e, lambda_e, xyz, a.disp, a.velo, a.acc,
                          a.rot})
    require_device(p, "a vector argument of the assembly");
  DFM2H_REQUIRE(M->np == A->nblk, DFM2H_ERR_INVALID, "mesh node count != matrix block rows");
  const size_t n = A->pat.ncrs + A->nblk;
  if (!n) return;
  DBuf<int> missing;
  missing.alloc(1);
  missing.zero();
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
  a.tet = reinterpret_cast<const uint4 *>(M->tet.p);
  a.elsup_ind = M->elsup_ind.p;
  a.elsup = M->elsup.p;
  a.missing = missing.p;
  const dim3 grid(ceil_div(n, 256)), block(256);
  if (walk_assembly(M)) {
    if (kind == kAsmNewmark)
      hipLaunchKernelGGL(k_assemble_tet_gather<kAsmNewmark>, grid, block, 0, stream(), a);
    else if (kind == kAsmStiffwarp)
      hipLaunchKernelGGL(k_assemble_tet_gather<kAsmStiffwarp>, grid, block, 0, stream(), a);
    else if (kind == kAsmBEuler)
      hipLaunchKernelGGL(k_assemble_tet_gather<kAsmBEuler>, grid, block, 0, stream(), a);
    else
      hipLaunchKernelGGL(k_assemble_tet_gather<kAsmStatic>, grid, block, 0, stream(), a);
    check_launch("k_assemble_tet_gather");
  } else {
    if (A->asm_mesh_serial != M->serial) {  // element lists for this mesh
      A->asm_mesh_serial = 0;
      DBuf<unsigned> cnt;
      cnt.alloc(n);
      hipLaunchKernelGGL(k_asm_lists, grid, block, 0, stream(), a, (const unsigned *)nullptr,
                         cnt.p);
      check_launch("k_asm_lists");
      std::vector<unsigned> ptr(n + 1, 0);
      DFM2H_HIP(hipMemcpyAsync(ptr.data() + 1, cnt.p, n * sizeof(unsigned),
                               hipMemcpyDeviceToHost, stream()));
      int h_miss = 0;
      DFM2H_HIP(hipMemcpyAsync(&h_miss, missing.p, sizeof(int), hipMemcpyDeviceToHost,
                               stream()));
      DFM2H_HIP(hipStreamSynchronize(stream()));
      size_t tot = 0;
      for (size_t q = 1; q <= n; ++q) {
        tot += ptr[q];
        DFM2H_REQUIRE(tot < (1ull << 32), DFM2H_ERR_INVALID, "too many element entries");
        ptr[q] = (unsigned)tot;
      }
      A->asm_ptr.upload(ptr);
      A->asm_ent.alloc(std::max<size_t>(tot, 1));
      hipLaunchKernelGGL(k_asm_lists, grid, block, 0, stream(), a, A->asm_ptr.p, A->asm_ent.p);
      check_launch("k_asm_lists");
      A->asm_missing = h_miss;
      A->asm_mesh_serial = M->serial;
    }
    A->asm_rec.alloc(std::max<size_t>(M->ntet * kTetRec, 1));
    if (M->ntet) {
      hipLaunchKernelGGL(k_tet_records, dim3(ceil_div(M->ntet, 256)), block, 0, stream(), a,
                         M->ntet, A->asm_rec.p);
      check_launch("k_tet_records");
    }
    const unsigned *pp = A->asm_ptr.p, *pe = A->asm_ent.p;
    const double *pr = A->asm_rec.p;
    const dim3 go(ceil_div(std::max<size_t>(A->pat.ncrs, 1), 256));
    const dim3 gd(ceil_div(A->nblk, 64 / kDiagLanes)), bd(64);
    const bool rows = M->max_elsup <= kRowElems;
#define DFM2H_ASM_OFFDIAG(K)                                                                  \
  do {                                                                                       \
    if (A->pat.ncrs && rows)                                                                 \
      hipLaunchKernelGGL(k_assemble_offdiag_rows<K>, gd, bd, 0, stream(), a, pp, pe, pr);    \
    else if (A->pat.ncrs)                                                                    \
      hipLaunchKernelGGL(k_assemble_offdiag_lists<K>, go, block, 0, stream(), a, pp, pe, pr); \
  } while (0)
#define DFM2H_ASM_LISTS(K)                                                                   \
  do {                                                                                       \
    DFM2H_ASM_OFFDIAG(K);                                                                    \
    hipLaunchKernelGGL(k_assemble_diag_lists<K>, gd, bd, 0, stream(), a, pp, pe, pr);        \
  } while (0)
    if (kind == kAsmNewmark) {
      DFM2H_ASM_OFFDIAG(kAsmNewmark);
      hipLaunchKernelGGL(k_assemble_diag_newmark, gd, bd, 0, stream(), a, pp, pe, pr);
    }
    else if (kind == kAsmStiffwarp) {
      DFM2H_ASM_OFFDIAG(kAsmStiffwarp);
      hipLaunchKernelGGL(k_assemble_diag_stiffwarp, gd, bd, 0, stream(), a, pp, pe, pr);
    }
    else if (kind == kAsmBEuler)
      DFM2H_ASM_LISTS(kAsmBEuler);
    else
      DFM2H_ASM_LISTS(kAsmStatic);
#undef DFM2H_ASM_LISTS
#undef DFM2H_ASM_OFFDIAG
    check_launch("k_assemble_lists");
  }
  int h_missing = 0;
  DFM2H_HIP(hipMemcpyAsync(&h_missing, missing.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  DFM2H_HIP(hipStreamSynchronize(stream()));
  if (!walk_assembly(M) && A->asm_missing) h_missing = 1;  // found when listing
  A->version++;
  DFM2H_REQUIRE(!h_missing, DFM2H_ERR_PATTERN,
                "an element edge has no block in the matrix pattern");
}

// XPlusAYBZ (vecxitrsol.cpp:312-325) and, with W, XPlusAYBZCW (:327-342):
// X += alpha Y + beta Z (+ gamma W) on the free entries, same association.
template <bool CW>
__global__ void k_xplusaybzcw(double *X, size_t n, const int *__restrict__ flag, double alpha,
                              const double *__restrict__ Y, double beta,
                              const double *__restrict__ Z, double gamma,
                              const double *__restrict__ W) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || flag[i] != 0) return;
  if (CW)
    X[i] += alpha * Y[i] + beta * Z[i] + gamma * W[i];
  else
    X[i] += alpha * Y[i] + beta * Z[i];
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
    {
      std::lock_guard<std::mutex> lk(g_live_mu);
      g_live.insert(A);
    }
    *out = A;
  });
}

int dfm2h_bsr_destroy(dfm2h_bsr A) {
  return guard(__func__, [&] {
    {
      std::lock_guard<std::mutex> lk(g_live_mu);
      g_live.erase(A);
    }
    delete A;
  });
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
    A->pattern_serial = next_serial();
    A->mg_mark.assign(A->nblk, ~0u);
    A->version++;
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_bsr_set_zero(dfm2h_bsr A) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    A->mg_dst.clear();  // staged additions are overwritten anyway
    A->mg_val.clear();
    A->vcrs.zero();
    A->vdia.zero();
    A->version++;
  });
}

int dfm2h_bsr_upload(dfm2h_bsr A, const double *vcrs, const double *vdia) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    flush_merge(A);
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
    flush_merge(A);
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
    if (vcrs || vdia) flush_merge(A);  // the device arrays hold every staged addition
    if (nblk) *nblk = A->nblk;
    if (ncrs) *ncrs = A->pat.ncrs;
    if (vcrs) *vcrs = A->vcrs.p;
    if (vdia) *vdia = A->vdia.p;
  });
}

int dfm2h_bsr_values_changed(dfm2h_bsr A) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    A->version++;
  });
}

// ls_block_sparse.h:260-310: row by row, the row's slots are marked by
// column (the reference's merge_buffer), then each column's block is added
// -- emat[irow][irow] when the column is the row itself, into the diagonal
// block -- and the marks are cleared.  A missing slot ends the call with
// the blocks before it staged, as the reference returns false there.
int dfm2h_bsr_merge(dfm2h_bsr A, unsigned nrow, const unsigned *ip_row, unsigned ncol,
                    const unsigned *ip_col, const double *emat) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A, DFM2H_ERR_INVALID, "null matrix handle");
    DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_STATE, "SetPattern has not been called");
    DFM2H_REQUIRE((ip_row && ip_col && emat) || nrow == 0 || ncol == 0, DFM2H_ERR_INVALID,
                  "null argument");
    for (unsigned r = 0; r < nrow; ++r)
      DFM2H_REQUIRE(ip_row[r] < A->nblk, DFM2H_ERR_INVALID, "row block index out of range");
    for (unsigned c = 0; c < ncol; ++c)
      DFM2H_REQUIRE(ip_col[c] < A->nblk, DFM2H_ERR_INVALID, "column block index out of range");
    if (A->mg_dst.size() + (size_t)nrow * ncol > kMergeChunk) flush_merge(A);
    const Pattern &pt = A->pat;
    std::vector<unsigned> &mark = A->mg_mark;
    for (unsigned r = 0; r < nrow; ++r) {
      const unsigned i = ip_row[r];
      for (unsigned q = pt.colind[i]; q < pt.colind[i + 1]; ++q) mark[pt.rowptr[q]] = q;
      bool missing = false;
      for (unsigned c = 0; c < ncol && !missing; ++c) {
        const unsigned j = ip_col[c];
        unsigned dst;
        const double *in;
        if (i == j) {
          DFM2H_REQUIRE(r < ncol, DFM2H_ERR_INVALID,
                        "a diagonal block takes emat[irow][irow], which needs irow < ncol");
          dst = (unsigned)pt.ncrs + i;
          in = emat + 9 * ((size_t)r * ncol + r);
        } else if (mark[j] != ~0u) {
          dst = mark[j];
          in = emat + 9 * ((size_t)r * ncol + c);
        } else {
          missing = true;
          break;
        }
        A->mg_dst.push_back(dst);
        A->mg_val.insert(A->mg_val.end(), in, in + 9);
      }
      for (unsigned q = pt.colind[i]; q < pt.colind[i + 1]; ++q) mark[pt.rowptr[q]] = ~0u;
      DFM2H_REQUIRE(!missing, DFM2H_ERR_PATTERN, "a block of the element is not in the pattern");
    }
  });
}

int dfm2h_bsr_matvec(dfm2h_bsr A, double *y, double alpha, const double *x, double beta) {
  return guard(__func__, [&] {
    need_pattern(A);
    DFM2H_REQUIRE(y && x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_REQUIRE(y != x, DFM2H_ERR_INVALID, "y and x must not alias");
    DFM2H_DEVICE_ARG(y);
    DFM2H_DEVICE_ARG(x);
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
    DFM2H_DEVICE_ARG(flag);
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
    DFM2H_DEVICE_ARG(b);
    DFM2H_DEVICE_ARG(flag);
    hipLaunchKernelGGL(k_set_rhs_zero, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), b, flag, n,
                       nz);
    check_launch("k_set_rhs_zero");
  });
}

int dfm2h_xplusay(double *X, size_t n, const int *flag, double alpha, const double *Y) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE((X && flag && Y) || n == 0, DFM2H_ERR_INVALID, "null vector");
    if (!n) return;
    DFM2H_DEVICE_ARG(X);
    DFM2H_DEVICE_ARG(flag);
    DFM2H_DEVICE_ARG(Y);
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
    AsmArgs a{};
    a.disp = disp;
    run_assembly(kAsmStatic, A, vec_b, myu_e, lambda_e, myu, lambda, rho, g, xyz, M, a);
  });
}

int dfm2h_merge_linsys_solidlinear_newmark_tet(dfm2h_bsr A, double *vec_b, const double *myu_e,
                                               const double *lambda_e, double myu, double lambda,
                                               double rho, const double *g, double dt,
                                               double gamma_newmark, double beta_newmark,
                                               const double *xyz, dfm2h_tetmesh M,
                                               const double *disp, const double *velo,
                                               const double *acc) {
  (void)gamma_newmark;  // unused by the element matrix (fem_solidlinear.cpp:447)
  return guard(__func__, [&] {
    DFM2H_REQUIRE(dt > 0.0, DFM2H_ERR_INVALID, "time step must be positive");
    AsmArgs a{};
    a.disp = disp;
    a.velo = velo;
    a.acc = acc;
    a.dt = dt;
    a.beta_nb = beta_newmark;
    run_assembly(kAsmNewmark, A, vec_b, myu_e, lambda_e, myu, lambda, rho, g, xyz, M, a);
  });
}

int dfm2h_merge_linsys_solidlinear_beuler_tet(dfm2h_bsr A, double *vec_b, const double *myu_e,
                                              const double *lambda_e, double myu, double lambda,
                                              double rho, const double *g, double dt,
                                              const double *xyz, dfm2h_tetmesh M,
                                              const double *disp, const double *velo) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(dt > 0.0, DFM2H_ERR_INVALID, "time step must be positive");
    AsmArgs a{};
    a.disp = disp;
    a.velo = velo;
    a.dt = dt;
    run_assembly(kAsmBEuler, A, vec_b, myu_e, lambda_e, myu, lambda, rho, g, xyz, M, a);
  });
}

int dfm2h_merge_linsys_solidstiffwarp_beuler_tet(dfm2h_bsr A, double *vec_b,
                                                  const double *myu_e, const double *lambda_e,
                                                  double myu, double lambda, double rho,
                                                  const double *g, double dt, const double *xyz,
                                                  dfm2h_tetmesh M, const double *disp,
                                                  const double *velo, const double *rot) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(dt > 0.0, DFM2H_ERR_INVALID, "time step must be positive");
    DFM2H_REQUIRE(rot, DFM2H_ERR_INVALID, "null rotation array");
    AsmArgs a{};
    a.disp = disp;
    a.velo = velo;
    a.rot = rot;
    a.dt = dt;
    run_assembly(kAsmStiffwarp, A, vec_b, myu_e, lambda_e, myu, lambda, rho, g, xyz, M, a);
  });
}

int dfm2h_xplusaybz(double *X, size_t n, const int *flag, double alpha, const double *Y,
                    double beta, const double *Z) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE((X && flag && Y && Z) || n == 0, DFM2H_ERR_INVALID, "null vector");
    if (!n) return;
    for (const void *p : {(const void *)X, (const void *)flag, (const void *)Y, (const void *)Z})
      require_device(p, "a vector argument");
    hipLaunchKernelGGL(k_xplusaybzcw<false>, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), X,
                       n, flag, alpha, Y, beta, Z, 0.0, (const double *)nullptr);
    check_launch("k_xplusaybz");
  });
}

int dfm2h_xplusaybzcw(double *X, size_t n, const int *flag, double alpha, const double *Y,
                      double beta, const double *Z, double gamma, const double *W) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE((X && flag && Y && Z && W) || n == 0, DFM2H_ERR_INVALID, "null vector");
    if (!n) return;
    for (const void *p : {(const void *)X, (const void *)flag, (const void *)Y, (const void *)Z,
                          (const void *)W})
      require_device(p, "a vector argument");
    hipLaunchKernelGGL(k_xplusaybzcw<true>, dim3(ceil_div(n, 256)), dim3(256), 0, stream(), X,
                       n, flag, alpha, Y, beta, Z, gamma, W);
    check_launch("k_xplusaybzcw");
  });
}

}  // extern "C"
