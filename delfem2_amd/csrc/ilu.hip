// ilu.hip -- CPreconditionerILU<double> ILU(0) on the device: symbolic set-up
// with level scheduling (host), value copy, level-parallel numeric
// factorisation and level-parallel triangular sweeps.
#include "handles.hpp"

#include <algorithm>
#include <numeric>

namespace dfm2h {

// CopyValue (ls_ilu_block_sparse.cpp:1009-1047): ILU slot a2l[s] receives A
// slot s; slots of the ILU pattern that A lacks stay 0.
__global__ void k_copy_value(size_t ncrs_a, const int *__restrict__ a2l,
                             const double *__restrict__ a_vcrs, double *__restrict__ l_vcrs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncrs_a * 9) return;
  const int d = a2l[t / 9];
  if (d >= 0) l_vcrs[(size_t)d * 9 + t % 9] = a_vcrs[t];
}

__global__ void k_zero(double *p, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// ---------------------------------------------------------------------------
// ILU(0) numeric factorisation of one level (ls_ilu_block_sparse.cpp:286-356),
// one thread per row, the reference's IKJ loop verbatim in arithmetic:
//   for k in L(i) ascending, for j in U(k) ascending, (i,j) in pattern:
//     A_ij -= A_ik * U_kj            (or D_i when j == i)
//   D_i <- inverse(D_i) unless |det| <= 1e-30 (counted and marked in
//          sing[i], left as is)
//   U_ij <- D_i * U_ij   (skipped with scale_u = 0: the state in which the
//          reference returns at its 11th singular pivot)
// Row i's columns are sorted, so the slot of j in row i is found by a
// pointer that only moves forward (j ascends for a fixed k).

__device__ inline double det3(const double *a) {
  return +a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5]
         - a[0] * a[7] * a[5] - a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
}

__global__ void __launch_bounds__(64)
k_ilu_decompose_level(const unsigned *__restrict__ rows, unsigned nrows,
                      const unsigned *__restrict__ colind, const unsigned *__restrict__ rowptr,
                      const unsigned *__restrict__ diaind, double *vcrs, double *vdia,
                      int *nsing, unsigned char *sing, int scale_u) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows) return;
  const unsigned i = rows[t];
  const unsigned i0 = colind[i], i1 = colind[i + 1], idia = diaind[i];
  for (unsigned ik = i0; ik < idia; ++ik) {
    const unsigned k = rowptr[ik];
    double vik[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) vik[q] = vcrs[(size_t)ik * 9 + q];
    unsigned pos = ik + 1;
    const unsigned k1 = colind[k + 1];
    for (unsigned kj = diaind[k]; kj < k1; ++kj) {
      const unsigned j = rowptr[kj];
      double *vij;
      if (j == i) {
        vij = vdia + (size_t)i * 9;
      } else {
        while (pos < i1 && rowptr[pos] < j) ++pos;
        if (pos >= i1 || rowptr[pos] != j) continue;
        vij = vcrs + (size_t)pos * 9;
      }
      const double *vkj = vcrs + (size_t)kj * 9;
      double u[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) u[q] = vkj[q];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        vij[a * 3 + 0] -= vik[a * 3 + 0] * u[0] + vik[a * 3 + 1] * u[3] + vik[a * 3 + 2] * u[6];
        vij[a * 3 + 1] -= vik[a * 3 + 0] * u[1] + vik[a * 3 + 1] * u[4] + vik[a * 3 + 2] * u[7];
        vij[a * 3 + 2] -= vik[a * 3 + 0] * u[2] + vik[a * 3 + 1] * u[5] + vik[a * 3 + 2] * u[8];
      }
    }
  }
  double d[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) d[q] = vdia[(size_t)i * 9 + q];
  if (fabs(det3(d)) > 1.0e-30) {
    // CalcInvMat3 (ls_ilu_block_sparse.cpp:84-103)
    const double inv_det = 1.0 / det3(d);
    double a[9];
    a[0] = inv_det * (d[4] * d[8] - d[5] * d[7]);
    a[1] = inv_det * (d[2] * d[7] - d[1] * d[8]);
    a[2] = inv_det * (d[1] * d[5] - d[2] * d[4]);
    a[3] = inv_det * (d[5] * d[6] - d[3] * d[8]);
    a[4] = inv_det * (d[0] * d[8] - d[2] * d[6]);
    a[5] = inv_det * (d[2] * d[3] - d[0] * d[5]);
    a[6] = inv_det * (d[3] * d[7] - d[4] * d[6]);
    a[7] = inv_det * (d[1] * d[6] - d[0] * d[7]);
    a[8] = inv_det * (d[0] * d[4] - d[1] * d[3]);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      d[q] = a[q];
      vdia[(size_t)i * 9 + q] = a[q];
    }
  } else {
    atomicAdd(nsing, 1);
    sing[i] = 1;
  }
  for (unsigned s = idia; scale_u && s < i1; ++s) {
    double *v = vcrs + (size_t)s * 9;
    double tb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) tb[q] = v[q];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v[a * 3 + 0] = d[a * 3 + 0] * tb[0] + d[a * 3 + 1] * tb[3] + d[a * 3 + 2] * tb[6];
      v[a * 3 + 1] = d[a * 3 + 0] * tb[1] + d[a * 3 + 1] * tb[4] + d[a * 3 + 2] * tb[7];
      v[a * 3 + 2] = d[a * 3 + 0] * tb[2] + d[a * 3 + 1] * tb[5] + d[a * 3 + 2] * tb[8];
    }
  }
}

// The same factorisation of one level with one lane per block of a row: G
// lanes per row (G a power of two, more than the longest row's block count,
// at most 64).  Lane t of row i's group holds block t of the row in CRS
// order with the diagonal between the lower and the upper part: t < nL the
// lower slot i0 + t, t == nL the diagonal, t > nL the upper slot i0 + t - 1.
// Step s takes the row's s-th lower block A_ik -- final by then, since every
// update it receives comes from a smaller k, i.e. an earlier step -- from
// lane s, and every lane whose column j lies in U(k) applies
// A_ij -= A_ik U_kj.  For each destination the updates come in ascending k
// with the products and association of k_ilu_decompose_level, and D^-1 and
// the U scaling are formed with its expressions, so both kernels give the
// same bits; this one keeps a row's dependent loads to about two memory
// round trips per lower block instead of one per (k, j) pair.
template <int G>
__global__ void __launch_bounds__(64)
k_ilu_decompose_lanes(const unsigned *__restrict__ rows, unsigned nrows,
                      const unsigned *__restrict__ colind, const unsigned *__restrict__ rowptr,
                      const unsigned *__restrict__ diaind, double *vcrs, double *vdia,
                      int *nsing, unsigned char *sing, int scale_u) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "G: power of two <= 64");
  const unsigned lane = threadIdx.x;  // blockDim.x == 64: one wavefront
  const unsigned gb = lane & ~(G - 1u), t = lane & (G - 1u);
  const unsigned ri = blockIdx.x * (64u / G) + lane / G;
  const bool live = ri < nrows;
  const unsigned i = live ? rows[ri] : 0u;
  unsigned i0 = 0, i1 = 0, idia = 0;
  if (live) {
    i0 = colind[i];
    i1 = colind[i + 1];
    idia = diaind[i];
  }
  const unsigned nL = idia - i0;
  const bool has = live && t < i1 - i0 + 1;
  const bool dia = has && t == nL;
  const size_t slot = t < nL ? (size_t)i0 + t : (size_t)i0 + t - 1;
  const unsigned j = !has ? ~0u : dia ? i : rowptr[slot];
  double *const dst = dia ? vdia + (size_t)i * 9 : vcrs + slot * 9;
  double v[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) v[q] = has ? dst[q] : 0.0;
  // upper part [kd, k1) of row k for this lane's lower block k
  unsigned kd = 0, k1 = 0;
  if (has && t < nL) {
    kd = diaind[j];
    k1 = colind[j + 1];
  }
  unsigned smax = live ? nL : 0u;  // steps: the wavefront's longest lower part
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) smax = max(smax, (unsigned)__shfl_xor((int)smax, off, 64));
  // row k's upper columns, one per lane of the group (requested a step ahead)
  auto upper_col = [&](unsigned s, unsigned &skd) -> unsigned {
    skd = (unsigned)__shfl((int)kd, (int)(gb + s), 64);
    const unsigned sk1 = (unsigned)__shfl((int)k1, (int)(gb + s), 64);
    return (live && s < nL && skd + t < sk1) ? rowptr[skd + t] : ~0u;
  };
  unsigned kd_next = 0;
  unsigned c_next = smax > 0 ? upper_col(0, kd_next) : ~0u;
  for (unsigned s = 0; s < smax; ++s) {
    const unsigned c = c_next, skd = kd_next;
    if (s + 1 < smax) c_next = upper_col(s + 1, kd_next);
    double m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = __shfl(v[q], (int)(gb + s), 64);
    unsigned kj = ~0u;
#pragma unroll
    for (unsigned u = 0; u < G; ++u) {
      const unsigned cu = (unsigned)__shfl((int)c, (int)(gb + u), 64);
      if (cu == j && j != ~0u) kj = skd + u;
    }
    if (kj != ~0u) {  // only where s < nL: c is ~0u otherwise
      const double *vkj = vcrs + (size_t)kj * 9;
      double u[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) u[q] = vkj[q];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        v[a * 3 + 0] -= m[a * 3 + 0] * u[0] + m[a * 3 + 1] * u[3] + m[a * 3 + 2] * u[6];
        v[a * 3 + 1] -= m[a * 3 + 0] * u[1] + m[a * 3 + 1] * u[4] + m[a * 3 + 2] * u[7];
        v[a * 3 + 2] -= m[a * 3 + 0] * u[2] + m[a * 3 + 1] * u[5] + m[a * 3 + 2] * u[8];
      }
    }
  }
  if (dia) {
    if (fabs(det3(v)) > 1.0e-30) {  // CalcInvMat3 (ls_ilu_block_sparse.cpp:84-103)
      const double inv_det = 1.0 / det3(v);
      double a[9];
      a[0] = inv_det * (v[4] * v[8] - v[5] * v[7]);
      a[1] = inv_det * (v[2] * v[7] - v[1] * v[8]);
      a[2] = inv_det * (v[1] * v[5] - v[2] * v[4]);
      a[3] = inv_det * (v[5] * v[6] - v[3] * v[8]);
      a[4] = inv_det * (v[0] * v[8] - v[2] * v[6]);
      a[5] = inv_det * (v[2] * v[3] - v[0] * v[5]);
      a[6] = inv_det * (v[3] * v[7] - v[4] * v[6]);
      a[7] = inv_det * (v[1] * v[6] - v[0] * v[7]);
      a[8] = inv_det * (v[0] * v[4] - v[1] * v[3]);
#pragma unroll
      for (int q = 0; q < 9; ++q) v[q] = a[q];
    } else {
      atomicAdd(nsing, 1);
      sing[i] = 1;
    }
  }
  double d[9];  // D_i^-1 (or D_i itself when singular), from the diagonal lane
#pragma unroll
  for (int q = 0; q < 9; ++q) d[q] = __shfl(v[q], (int)(gb + (nL & (G - 1u))), 64);
  if (scale_u && has && t > nL) {
    double tb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) tb[q] = v[q];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v[a * 3 + 0] = d[a * 3 + 0] * tb[0] + d[a * 3 + 1] * tb[3] + d[a * 3 + 2] * tb[6];
      v[a * 3 + 1] = d[a * 3 + 0] * tb[1] + d[a * 3 + 1] * tb[4] + d[a * 3 + 2] * tb[7];
      v[a * 3 + 2] = d[a * 3 + 0] * tb[2] + d[a * 3 + 1] * tb[5] + d[a * 3 + 2] * tb[8];
    }
  }
  if (has) {
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = v[q];
  }
}

// lanes per row for k_ilu_decompose_lanes: the smallest power of two above
// the longest row's block count (diagonal included), 0 when that exceeds 64
static unsigned decompose_lanes(const std::vector<unsigned> &colind, unsigned nblk) {
  unsigned mx = 0;
  for (unsigned i = 0; i < nblk; ++i) mx = std::max(mx, colind[i + 1] - colind[i] + 1);
  unsigned g = 8;
  while (g <= mx) g *= 2;
  return g <= 64 ? g : 0u;
}

// A slot -> ILU slot by column lookup in the ILU row (the row2crs map of
// CopyValue), -1 where the ILU pattern lacks the column.
static std::vector<int> build_a2l(const dfm2h_ilu_s &P, const Pattern &A) {
  std::vector<int> a2l(A.ncrs, -1);
  std::vector<int> row2crs(P.nblk, -1);
  for (unsigned i = 0; i < P.nblk; ++i) {
    for (unsigned s = P.colind[i]; s < P.colind[i + 1]; ++s) row2crs[P.rowptr[s]] = (int)s;
    for (unsigned s = A.colind[i]; s < A.colind[i + 1]; ++s) a2l[s] = row2crs[A.rowptr[s]];
    for (unsigned s = P.colind[i]; s < P.colind[i + 1]; ++s) row2crs[P.rowptr[s]] = -1;
  }
  return a2l;
}

}  // namespace dfm2h

using namespace dfm2h;

// Builds the A -> ILU slot map for CopyValue against A; notes when it is the
// identity (ILU(0) on A's own pattern), where CopyValue is one device copy.
static void set_a2l(dfm2h_ilu P, dfm2h_bsr A) {
  const std::vector<int> a2l = build_a2l(*P, A->pat);
  bool id = a2l.size() == P->ncrs;
  for (size_t s = 0; id && s < a2l.size(); ++s) id = a2l[s] == (int)s;
  P->a2l_d.upload(a2l);
  P->a2l_identity = id;
  P->mapped_for = A;
  P->mapped_serial = A->pattern_serial;
}

static void need_ilu(dfm2h_ilu P) {
  DFM2H_REQUIRE(P, DFM2H_ERR_INVALID, "null preconditioner handle");
  DFM2H_REQUIRE(P->has_pattern, DFM2H_ERR_STATE, "SetPattern0 has not been called");
}

// Factorises the rows rows[0 .. nrows) (one dependency level, or one row),
// with or without the U scaling.
static void decompose_rows(dfm2h_ilu P, const unsigned *rows, unsigned nrows, int scale_u) {
  if (!nrows) return;
  const unsigned G = P->dec_lanes;
  if (G == 0) {
    hipLaunchKernelGGL(k_ilu_decompose_level, dim3(ceil_div(nrows, 64)), dim3(64), 0, stream(),
                       rows, nrows, P->colind_d.p, P->rowptr_d.p, P->diaind_d.p, P->vcrs.p,
                       P->vdia.p, P->nsing_d.p, P->sing_d.p, scale_u);
    return;
  }
  const dim3 grid(ceil_div((size_t)nrows * G, 64));
#define DFM2H_DEC_LANES(g)                                                                    \
  hipLaunchKernelGGL(k_ilu_decompose_lanes<g>, grid, dim3(64), 0, stream(), rows, nrows,     \
                     P->colind_d.p, P->rowptr_d.p, P->diaind_d.p, P->vcrs.p, P->vdia.p,      \
                     P->nsing_d.p, P->sing_d.p, scale_u)
  switch (G) {
    case 8: DFM2H_DEC_LANES(8); break;
    case 16: DFM2H_DEC_LANES(16); break;
    case 32: DFM2H_DEC_LANES(32); break;
    default: DFM2H_DEC_LANES(64); break;
  }
#undef DFM2H_DEC_LANES
}

// The reference factorises rows in ascending order and returns at its 11th
// singular pivot (ls_ilu_block_sparse.cpp:331-333): rows before that row i*
// factorised, row i* eliminated but its U part not scaled, later rows as
// CopyValue left them.  The level-scheduled factorisation ran every row;
// rows before i* hold the reference's values already (they depend on
// earlier rows only), so rows from i* on are copied again from the matrix
// CopyValue read -- when that matrix still exists unchanged -- and row i* is
// redone without the U scaling.  Without that matrix the rows after i* stay
// factorised.
static void stop_at_eleventh_singular(dfm2h_ilu P) {
  std::vector<unsigned char> sing(P->nblk);
  DFM2H_HIP(hipMemcpyAsync(sing.data(), P->sing_d.p, P->nblk, hipMemcpyDeviceToHost, stream()));
  DFM2H_HIP(hipStreamSynchronize(stream()));
  unsigned stop = P->nblk, cnt = 0;
  for (unsigned i = 0; i < P->nblk && stop == P->nblk; ++i)
    if (sing[i] && ++cnt == 11) stop = i;
  const dfm2h_bsr_s *A = P->cv_from;
  if (stop == P->nblk || !A || !bsr_alive(A) || A->pattern_serial != P->cv_serial ||
      A->version != P->cv_version || P->mapped_for != A || P->mapped_serial != A->pattern_serial)
    return;
  const size_t l0 = P->colind[stop], a0 = A->pat.colind[stop];
  if (P->ncrs > l0)
    hipLaunchKernelGGL(k_zero, dim3(ceil_div((P->ncrs - l0) * 9, 256)), dim3(256), 0, stream(),
                       P->vcrs.p + l0 * 9, (P->ncrs - l0) * 9);
  if (A->pat.ncrs > a0)
    hipLaunchKernelGGL(k_copy_value, dim3(ceil_div((A->pat.ncrs - a0) * 9, 256)), dim3(256), 0,
                       stream(), A->pat.ncrs - a0, P->a2l_d.p + a0, A->vcrs.p + a0 * 9,
                       P->vcrs.p);
  DFM2H_HIP(hipMemcpyAsync(P->vdia.p + (size_t)stop * 9, A->vdia.p + (size_t)stop * 9,
                           sizeof(double) * 9 * (P->nblk - stop), hipMemcpyDeviceToDevice,
                           stream()));
  DBuf<unsigned> row;
  row.upload(&stop, 1);
  decompose_rows(P, row.p, 1, 0);
  check_launch("ILU stop row");
  DFM2H_HIP(hipStreamSynchronize(stream()));
}

namespace dfm2h {

void ilu_forward_levels_double(dfm2h_ilu P, const double *src, double *vec, hipStream_t s,
                               const int *done) {
  ensure_packed(P->fplan, P->f64, P->vcrs.p, P->vdia.p, P->version, s);
  for (size_t l = 0; l + 1 < P->fplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->fplan.group_slices(l);
    hipLaunchKernelGGL((k_sliced_fwd<double, 3>), dim3(slice_grid(nsl)), dim3(kPackBlock), 0, s,
                       P->fplan.group_ptr[l], nsl, view(P->fplan), P->fplan.gview[l],
                       P->f64.val.p, P->f64.dia.p, src, vec, done);
  }
  check_launch("k_sliced_fwd");
}

void ilu_backward_levels_double(dfm2h_ilu P, double *vec, hipStream_t s, const int *done) {
  ensure_packed(P->bplan, P->b64, P->vcrs.p, nullptr, P->version, s);
  for (size_t l = 0; l + 1 < P->bplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->bplan.group_slices(l);
    hipLaunchKernelGGL((k_sliced_bwd<double, 3>), dim3(slice_grid(nsl)), dim3(kPackBlock), 0, s,
                       P->bplan.group_ptr[l], nsl, view(P->bplan), P->bplan.gview[l],
                       P->b64.val.p, vec, done);
  }
  check_launch("k_sliced_bwd");
}

}  // namespace dfm2h

extern "C" {

int dfm2h_ilu_create(dfm2h_ilu *out) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    *out = new dfm2h_ilu_s;
  });
}

int dfm2h_ilu_destroy(dfm2h_ilu P) {
  return guard(__func__, [&] { delete P; });
}

// Everything after the ILU pattern (colind / rowptr / diaind on the host) is
// known: device copies, level schedules, sliced sweep plans, the A -> ILU
// slot map.
static void finish_pattern(dfm2h_ilu P, dfm2h_bsr A) {
  P->ncrs = P->colind[P->nblk];
  P->colind_d.upload(P->colind);
  P->rowptr_d.upload(P->rowptr);
  P->diaind_d.upload(P->diaind);
  P->dec_lanes = decompose_lanes(P->colind, P->nblk);
  P->vcrs.alloc(P->ncrs * 9);
  P->vdia.alloc((size_t)P->nblk * 9);
  std::vector<unsigned> rows;
  {
    std::vector<unsigned> s0(P->nblk), s1(P->nblk);
    build_levels(P->nblk, P->colind, P->rowptr, P->diaind, true, P->levptr_f, rows);
    P->levrows_f.upload(rows);
    P->flev_ranges = true;  // level l is rows [levptr_f[l], levptr_f[l+1]) in order
    for (unsigned k = 0; P->flev_ranges && k < P->nblk; ++k) P->flev_ranges = rows[k] == k;
    std::vector<std::vector<unsigned>> groups;
    for (size_t l = 0; l + 1 < P->levptr_f.size(); ++l)
      groups.emplace_back(rows.begin() + P->levptr_f[l], rows.begin() + P->levptr_f[l + 1]);
    for (unsigned i = 0; i < P->nblk; ++i) {
      s0[i] = P->colind[i];
      s1[i] = P->diaind[i];
    }
    P->fplan.build(groups, s0, s1, P->rowptr, true);
    {
      std::vector<unsigned> um;
      P->flev_u0.assign(groups.size(), 0);
      for (size_t l = 0; l < groups.size(); ++l)
        for (size_t b = 0; b < groups[l].size(); b += kSliceRows) {
          unsigned m = 0;
          for (size_t k = b; k < std::min(groups[l].size(), b + kSliceRows); ++k) {
            const unsigned i = groups[l][k];
            if (P->diaind[i] == P->colind[i + 1]) m |= 1u << (k - b);
          }
          um.push_back(m);
          if (m) P->flev_u0[l] = 1;
        }
      P->fumask.upload(um);
    }
    build_levels(P->nblk, P->colind, P->rowptr, P->diaind, false, P->levptr_b, rows);
    groups.clear();
    for (size_t l = 0; l + 1 < P->levptr_b.size(); ++l)
      groups.emplace_back(rows.begin() + P->levptr_b[l], rows.begin() + P->levptr_b[l + 1]);
    for (unsigned i = 0; i < P->nblk; ++i) {
      s0[i] = P->diaind[i];
      s1[i] = P->colind[i + 1];
    }
    P->bplan.build(groups, s0, s1, P->rowptr, true);
  }
  set_a2l(P, A);
  P->nsing_d.alloc(1);
  P->sing_d.alloc(std::max<size_t>(P->nblk, 1));
  P->has_pattern = true;
  P->pattern_serial = next_serial();
  P->decomposed = false;
  P->version++;
  DFM2H_HIP(hipStreamSynchronize(stream()));
}

// first column > i in the sorted row (or the row end)
static void fill_diaind(dfm2h_ilu P) {
  P->diaind.assign(P->nblk, 0);
  for (unsigned i = 0; i < P->nblk; ++i) {
    P->diaind[i] = P->colind[i + 1];
    for (unsigned s = P->colind[i]; s < P->colind[i + 1]; ++s)
      if (P->rowptr[s] > i) {
        P->diaind[i] = s;
        break;
      }
  }
}

// Symbolic ILU(k) (ls_ilu_block_sparse.cpp:849-999), lev_fill >= 1 or -1
// (every fill).  Row i starts from A's sorted columns at level 0; its
// entries k < i are taken in ascending order (fills added on the way
// included), skipping those with level + 1 > lev_fill; each upper entry j
// of row k (level + 1 <= lev_fill), j != i, not yet in row i, is inserted
// with level max(lev_ik, lev_kj) + 1.  An entry found again keeps its
// first level, as in the reference.
static void symbolic_iluk(dfm2h_ilu P, const Pattern &pa, int lev_fill) {
  const unsigned n = pa.nblk;
  std::vector<unsigned> col, lev;  // all finished rows
  std::vector<unsigned> dia(n);
  P->colind.assign(n + 1, 0);
  std::vector<std::pair<unsigned, unsigned>> row;  // (column, level), sorted
  auto allowed = [&](unsigned l) { return lev_fill < 0 || (int)l + 1 <= lev_fill; };
  for (unsigned i = 0; i < n; ++i) {
    row.clear();
    for (unsigned s = pa.colind[i]; s < pa.colind[i + 1]; ++s) row.emplace_back(pa.rowptr[s], 0u);
    std::sort(row.begin(), row.end());
    for (size_t t = 0; t < row.size(); ++t) {
      const unsigned k = row[t].first, lik = row[t].second;
      if (!allowed(lik)) continue;
      if (k >= i) break;
      for (unsigned kj = dia[k]; kj < P->colind[k + 1]; ++kj) {
        const unsigned j = col[kj], lkj = lev[kj];
        if (!allowed(lkj) || j == i) continue;
        auto it = std::lower_bound(row.begin() + t + 1, row.end(), std::make_pair(j, 0u),
                                   [](const auto &a, const auto &b) { return a.first < b.first; });
        if (it != row.end() && it->first == j) continue;
        row.insert(it, std::make_pair(j, std::max(lik, lkj) + 1));
      }
    }
    const size_t base = col.size();
    for (const auto &e : row) {
      col.push_back(e.first);
      lev.push_back(e.second);
    }
    DFM2H_REQUIRE(col.size() < (1ull << 32), DFM2H_ERR_INVALID, "ILU(k) pattern too large");
    P->colind[i + 1] = (unsigned)col.size();
    dia[i] = P->colind[i + 1];
    for (size_t t = 0; t < row.size(); ++t)
      if (row[t].first > i) {
        dia[i] = (unsigned)(base + t);
        break;
      }
  }
  P->rowptr = std::move(col);
  P->diaind = std::move(dia);
}

// ls_ilu_block_sparse.cpp:1057-1090: A's pattern, rows sorted
static void pattern0(dfm2h_ilu P, dfm2h_bsr A) {
  const Pattern &pa = A->pat;
  P->nblk = pa.nblk;
  P->colind = pa.colind;
  P->rowptr = pa.rowptr;
  for (unsigned i = 0; i < P->nblk; ++i)
    std::sort(P->rowptr.begin() + P->colind[i], P->rowptr.begin() + P->colind[i + 1]);
  fill_diaind(P);
}

int dfm2h_ilu_set_pattern0(dfm2h_ilu P, dfm2h_bsr A) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P && A, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_STATE, "matrix has no pattern");
    pattern0(P, A);
    finish_pattern(P, A);
  });
}

int dfm2h_ilu_info(dfm2h_ilu P, size_t *nblk, size_t *ncrs, unsigned *colind) {
  return guard(__func__, [&] {
    need_ilu(P);
    if (nblk) *nblk = P->nblk;
    if (ncrs) *ncrs = P->ncrs;
    if (colind) std::copy(P->colind.begin(), P->colind.end(), colind);
  });
}

int dfm2h_ilu_set_pattern_k(dfm2h_ilu P, dfm2h_bsr A, int lev_fill) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P && A, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(A->has_pattern, DFM2H_ERR_STATE, "matrix has no pattern");
    DFM2H_REQUIRE(lev_fill >= -1, DFM2H_ERR_INVALID, "lev_fill must be >= -1");
    if (lev_fill == 0) {
      pattern0(P, A);
    } else {
      P->nblk = A->pat.nblk;
      symbolic_iluk(P, A->pat, lev_fill);
    }
    finish_pattern(P, A);
  });
}

int dfm2h_ilu_copy_value(dfm2h_ilu P, dfm2h_bsr A) {
  return guard(__func__, [&] {
    need_ilu(P);
    DFM2H_REQUIRE(A && A->has_pattern, DFM2H_ERR_STATE, "matrix has no pattern");
    DFM2H_REQUIRE(A->nblk == P->nblk, DFM2H_ERR_INVALID, "size mismatch");
    flush_merge(A);
    // the map is rebuilt for another matrix, or for a matrix created where
    // the mapped one was (serials are process-unique)
    if (P->mapped_for != A || P->mapped_serial != A->pattern_serial) set_a2l(P, A);
    if (P->a2l_identity) {  // same pattern, same slots: a plain copy
      if (P->ncrs)
        DFM2H_HIP(hipMemcpyAsync(P->vcrs.p, A->vcrs.p, sizeof(double) * 9 * P->ncrs,
                                 hipMemcpyDeviceToDevice, stream()));
    } else {
      if (P->ncrs) {
        hipLaunchKernelGGL(k_zero, dim3(ceil_div(P->ncrs * 9, 256)), dim3(256), 0, stream(),
                           P->vcrs.p, P->ncrs * 9);
        check_launch("k_zero");
      }
      if (A->pat.ncrs) {
        hipLaunchKernelGGL(k_copy_value, dim3(ceil_div(A->pat.ncrs * 9, 256)), dim3(256), 0,
                           stream(), A->pat.ncrs, P->a2l_d.p, A->vcrs.p, P->vcrs.p);
        check_launch("k_copy_value");
      }
    }
    if (P->nblk)
      DFM2H_HIP(hipMemcpyAsync(P->vdia.p, A->vdia.p, sizeof(double) * 9 * P->nblk,
                               hipMemcpyDeviceToDevice, stream()));
    P->cv_from = A;
    P->cv_serial = A->pattern_serial;
    P->cv_version = A->version;
    P->decomposed = false;
    P->version++;
  });
}

int dfm2h_ilu_decompose(dfm2h_ilu P, int *ok, int *n_singular) {
  return guard(__func__, [&] {
    need_ilu(P);
    P->nsing_d.zero();
    P->sing_d.zero();
    for (size_t l = 0; l + 1 < P->levptr_f.size(); ++l)
      decompose_rows(P, P->levrows_f.p + P->levptr_f[l], P->levptr_f[l + 1] - P->levptr_f[l], 1);
    check_launch("k_ilu_decompose");
    int ns = 0;
    DFM2H_HIP(hipMemcpyAsync(&ns, P->nsing_d.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
    if (ns > 10) stop_at_eleventh_singular(P);
    if (ok) *ok = ns > 10 ? 0 : 1;
    // the reference stops counting where it stops factorising
    if (n_singular) *n_singular = ns > 10 ? 11 : ns;
    P->decomposed = true;
    P->version++;
  });
}

int dfm2h_ilu_forward(dfm2h_ilu P, double *vec) {
  return guard(__func__, [&] {
    need_ilu(P);
    DFM2H_REQUIRE(vec, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(vec);
    ilu_forward_levels_double(P, vec, vec, stream());
  });
}

int dfm2h_ilu_backward(dfm2h_ilu P, double *vec) {
  return guard(__func__, [&] {
    need_ilu(P);
    DFM2H_REQUIRE(vec, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(vec);
    ilu_backward_levels_double(P, vec, stream());
  });
}

int dfm2h_ilu_solve_precond(dfm2h_ilu P, double *vec) {
  return guard(__func__, [&] {
    need_ilu(P);
    DFM2H_REQUIRE(vec, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(vec);
    ilu_forward_levels_double(P, vec, vec, stream());
    ilu_backward_levels_double(P, vec, stream());
  });
}

int dfm2h_ilu_download(dfm2h_ilu P, unsigned *rowptr, unsigned *diaind, double *vcrs,
                       double *vdia) {
  return guard(__func__, [&] {
    need_ilu(P);
    if (rowptr) std::copy(P->rowptr.begin(), P->rowptr.end(), rowptr);
    if (diaind) std::copy(P->diaind.begin(), P->diaind.end(), diaind);
    if (vcrs) P->vcrs.download(vcrs, P->ncrs * 9);
    if (vdia) P->vdia.download(vdia, (size_t)P->nblk * 9);
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_ilu_levels(dfm2h_ilu P, unsigned *nf, unsigned *nb) {
  return guard(__func__, [&] {
    need_ilu(P);
    if (nf) *nf = (unsigned)(P->levptr_f.size() - 1);
    if (nb) *nb = (unsigned)(P->levptr_b.size() - 1);
  });
}

}  // extern "C"
