// batch.hip -- many independent tet systems solved together.
//
// The meshes of a batch are the diagonal blocks of one union system
// (A = diag(A_0 .. A_{n-1}), ILU(0) factors likewise), assembled and
// factorised with the ordinary entry points.  The solver runs one
// preconditioned CG per mesh, all in the same kernels:
//   * the packed plans are rebuilt with every slice inside one mesh: the
//     matrix plan groups rows by mesh, the sweep plans by (level, mesh), so
//     one sweep launch per dependency level serves every mesh;
//   * every reduction writes one partial per slice; a scalar kernel runs one
//     wavefront per mesh over that mesh's slices (a segmented, fixed-order
//     and therefore deterministic sum) and keeps alpha, beta, (r, Pr), the
//     iteration count and the convergence state per mesh;
//   * a wavefront whose mesh has converged returns at once, so a mesh stops
//     costing bandwidth the iteration after it converges;
//   * iterations are captured in a HIP graph (16 per graph); the host polls
//     the number of finished meshes one graph behind.
// FP64 and FP32-inner / FP64-reliable-update forms are both provided.  The
// launch shape (wavefronts per workgroup, slices per wavefront, blocks per
// load round) is chosen by a built-in autotuner per (DOF bucket,
// precision), cached in-process and optionally in a text file.
#include "handles.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

namespace dfm2h {
namespace {

constexpr int kVSB = 3;            // node stride of the batch vectors (caller layout)
constexpr int kMeshPerBlock = 4;   // scalar kernels: one wavefront per mesh
constexpr unsigned kChunk = 16;    // iterations per captured graph

struct MeshScal {
  double rPr, alpha, beta, inv_sq0, sq_b, rmax2, res;
  double best2;  // FP32 form: smallest FP64 |r|^2 of an update that at least halved |r|
  int it, done, converged, nan, upd, nupd, final_, it_upd, restart;
  int it_best;   // inner iteration of that update
  int stalled;   // kStallIters inner iterations since it: stopped, for the FP64 finish
  int pad;
};

struct BatchGlobal {
  double tol, delta2;
  int max_nitr, max_upd;
  unsigned ndone;
  int period;      // FP32 form: force a reliable update after this many iterations
  int keep_scale;  // FP64 form on a correction: the stop test stays relative to the
                   // |b|^2 a preceding mixed solve left in sq_b (set for no mesh: 0)
};

struct BPlan {
  const unsigned *off, *kmax, *row, *len, *col, *mesh;
};

template <typename T>
struct BArgs {
  unsigned nmesh, nlev_b, nlev_f;
  int any_iso;
  BPlan A, F, B;
  const T *a_val, *a_dia, *f_val, *f_dia, *b_val;
  const double *a64_val, *a64_dia;
  // fused forward level 0 / backward level 0 (matrix-slice masks, D^-1 of
  // the level-0 rows in matrix-slice layout; per forward slice: rows with no
  // upper blocks; per forward level: whether it has such rows)
  const unsigned *l0mask, *isomask, *fumask;
  const T *dinv0;
  const unsigned char *flev_u0;
  T *r, *x, *Pr, *p;
  double *x64;
  const double *b64;
  // per-slice partials: partA (p,Ap) / |b|^2, partA2 |r|^2, partB and partF
  // and partI (r,Pr), partE reliable-update drift
  double *partA, *partA2, *partB, *partF, *partI, *partE;
  MeshScal *M;
  BatchGlobal *G;
  const unsigned *a_mptr, *b_gptr, *f_gptr;
};

// Launch-shape variants the autotuner chooses from:
//   WPB  wavefronts (slices in flight) per workgroup
//   SPW  slices handled by one wavefront, one after the other
//   US   blocks per load round in the SpMV / residual (FP32, FP64)
//   UW   blocks per load round in the triangular sweeps
template <int WPB_, int SPW_, int US32_, int US64_, int UW_>
struct Var {
  static constexpr int WPB = WPB_, SPW = SPW_, UW = UW_;
  template <typename T>
  static constexpr int US() {
    return sizeof(T) == 8 ? US64_ : US32_;
  }
};
using V0 = Var<4, 1, 8, 4, 16>;
using V1 = Var<4, 1, 4, 2, 8>;
using V2 = Var<2, 1, 8, 4, 16>;
using V3 = Var<8, 1, 8, 4, 16>;
using V4 = Var<4, 2, 8, 4, 16>;
using V5 = Var<1, 1, 4, 4, 8>;
constexpr int kNumVariants = 6;
const char *const kVariantNames[kNumVariants] = {
    "wpb4_spw1_u8/4_w16", "wpb4_spw1_u4/2_w8", "wpb2_spw1_u8/4_w16",
    "wpb8_spw1_u8/4_w16", "wpb4_spw2_u8/4_w16", "wpb1_spw1_u4/4_w8"};

template <class V>
__device__ inline SliceLane lane_of(unsigned s) {
  SliceLane L;
  L.lane = threadIdx.x & (kWave - 1);
  L.s = s;
  L.c = L.lane / kSliceRows;
  L.r = L.lane - L.c * kSliceRows;
  return L;
}

// A slice's mesh, lane L's row, the row's block count and the slice's slot
// range, all loaded in one round (no load waits for another; a padding row
// has row ~0u and length 0).
struct RowMeta {
  unsigned mesh, i, len, kmax, off;
};

template <class P>
__device__ inline RowMeta row_meta(const P &pl, const SliceLane &L, unsigned s) {
  RowMeta R;
  R.mesh = pl.mesh[s];
  R.i = L.c < 3 ? pl.row[s * kSliceRows + L.r] : ~0u;
  R.len = L.c < 3 ? pl.len[s * kSliceRows + L.r] : 0u;
  R.kmax = pl.kmax[s];
  R.off = pl.off[s];
  return R;
}

// slices [slice0, slice0 + nsl): wavefront w of the grid takes SPW
// consecutive ones.  (Numbering the workgroups per XCD, as the
// single-system solver can, measured 3% slower here: the meshes converge at
// different iterations, and contiguous runs per XCD leave XCDs idle.)
template <class V>
__device__ inline unsigned wave_first(unsigned slice0) {
  return slice0 + (blockIdx.x * V::WPB + threadIdx.x / kWave) * V::SPW;
}

template <class V>
inline unsigned grid_of(unsigned nsl) {
  return std::max(1u, ceil_div(nsl, (size_t)V::WPB * V::SPW));
}

// ---------------------------------------------------------------------------
// slice kernels (every one walks the slices of one plan; a wavefront is
// uniform in slice, hence in mesh).  The iteration has the fused form of the
// single-system solver: forward level 0 (rows with no lower blocks) is
// computed by the r/x update from the row's new r components; backward level
// 0 (rows with no upper blocks) has no work, its (r, Pr) share is summed by
// the forward level that finalises those rows (partF) or, for rows with
// neither lower nor upper blocks, by the update (partI); the FP32 form's
// reliable update runs inside the next SpMV.

// r = b, x = 0 (and x64 = 0), partial |b|^2
template <typename T, class V>
__global__ void __launch_bounds__(V::WPB * kWave) kb_init(BArgs<T> a, unsigned nsl) {
  const unsigned s0 = wave_first<V>(0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= nsl) return;
    const SliceLane L = lane_of<V>(s);
    const unsigned i = L.c < 3 ? a.A.row[s * kSliceRows + L.r] : ~0u;
    double part = 0.0;
    if (i != ~0u) {
      const size_t e = kVSB * (size_t)i + L.c;
      const double v = a.b64[e];
      part = v * v;
      a.r[e] = (T)v;
      a.x[e] = (T)0;
      if (sizeof(T) == 4) a.x64[e] = 0.0;
    }
    part = wave_reduce_sum(part);
    if (L.lane == 0) a.partA[s] = part;
  }
}

// FP32 form, reliable update of row i: r_i = b_i - (A (x64 + x32))_i in FP64
// (FP64 values), stored into r; returns r_i^2 and adds (r_true - r_rec)_i^2
// to drift.
template <class V>
__device__ inline double batch_residual_row(const BArgs<float> &a, const SliceLane &L,
                                            unsigned s, unsigned i, double &drift) {
  const unsigned K = a.A.len[s * kSliceRows + L.r];
  const unsigned q0 = a.A.off[s];
  double acc = 0.0;
  for (unsigned k = 0; k < K; ++k) {
    const size_t q = q0 + k;
    const size_t j = kVSB * (size_t)a.A.col[q * kSliceRows + L.r];
    const Trip<double> v = load_trip(a.a64_val, q, L.lane);
    acc += v.a * (a.x64[j] + (double)a.x[j]) + v.b * (a.x64[j + 1] + (double)a.x[j + 1]) +
           v.c * (a.x64[j + 2] + (double)a.x[j + 2]);
  }
  const size_t ii = kVSB * (size_t)i;
  acc += dia_row_dot(a.a64_dia, L, a.x64[ii] + (double)a.x[ii],
                     a.x64[ii + 1] + (double)a.x[ii + 1], a.x64[ii + 2] + (double)a.x[ii + 2]);
  const double r = a.b64[ii + L.c] - acc;
  const double e = r - (double)a.r[ii + L.c];
  drift += e * e;
  a.r[ii + L.c] = (float)r;
  return r * r;
}

// Ap -> Pr, partial (p, Ap); FP32 form with the mesh's reliable update due:
// also r = b - A x (partA2: |r|^2, partE: drift)
template <typename T, class V>
__global__ void __launch_bounds__(V::WPB * kWave) kb_spmv(BArgs<T> a, unsigned nsl) {
  const unsigned s0 = wave_first<V>(0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= nsl) return;
    const SliceLane L = lane_of<V>(s);
    const RowMeta R = row_meta(a.A, L, s);
    const MeshScal *S = a.M + R.mesh;
    const int dn = ld_state(&S->done);  // requested together with upd
    const bool refine = sizeof(T) == 4 && ld_state(&S->upd);  // wavefront-uniform
    if (dn) continue;
    const unsigned i = R.i;
    double part = 0.0, part2 = 0.0, part3 = 0.0;
    if (i != ~0u) {
      // the row's own node and diagonal block first: off the gathers' chain
      const Trip<T> pi = load_node<T, kVSB>(a.p, i);
      const Trip<T> d = load_trip<T>(a.a_dia, s, L.lane);
      T acc;
      if constexpr (sizeof(T) == 4)  // FP32 copy: difference form, d = row-sum block
        acc = sliced_row_diff_sum<T, T, V::template US<T>(), kVSB>(
            (T)0, pi, L, R.len, R.len, R.kmax, R.off, a.A.col, a.a_val, a.p);
      else
        acc = sliced_row_sum<T, false, V::template US<T>(), kVSB>(
            (T)0, (T)1, L, R.len, R.kmax, R.off, a.A.col, a.a_val, a.p);
      acc += (T)d.a * pi.a + (T)d.b * pi.b + (T)d.c * pi.c;  // dia_row_dot's products
      const size_t e = kVSB * (size_t)i + L.c;
      a.Pr[e] = acc;
      part = (double)(L.c == 0 ? pi.a : L.c == 1 ? pi.b : pi.c) * (double)acc;
      if constexpr (sizeof(T) == 4)
        if (refine) part2 = batch_residual_row<V>(a, L, s, i, part3);
    }
    part = wave_reduce_sum(part);
    if (L.lane == 0) a.partA[s] = part;
    if (refine) {
      part2 = wave_reduce_sum(part2);
      part3 = wave_reduce_sum(part3);
      if (L.lane == 0) {
        a.partA2[s] = part2;
        a.partE[s] = part3;
      }
    }
  }
}

// r -= alpha Ap, x += alpha p (FP32 form after a reliable update: x64 +=
// x32, x32 = alpha p), partial |r|^2; then forward level 0 of the slice's
// level-0 rows (and (r, Pr) of its isolated rows)
template <typename T, class V>
__global__ void __launch_bounds__(V::WPB * kWave) kb_update(BArgs<T> a, unsigned nsl) {
  const unsigned s0 = wave_first<V>(0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= nsl) return;
    const SliceLane L = lane_of<V>(s);
    const unsigned mesh = a.A.mesh[s];
    const unsigned i = L.c < 3 ? a.A.row[s * kSliceRows + L.r] : ~0u;
    const MeshScal *S = a.M + mesh;
    // the state fields requested together (one memory round trip)
    const int dn = ld_state(&S->done);
    const double alpha = ld_state(&S->alpha);
    const bool fold = sizeof(T) == 4 && ld_state(&S->upd);
    if (dn) continue;
    double part = 0.0, srp = 0.0;
    T rn = 0;
    if (i != ~0u) {
      const size_t e = kVSB * (size_t)i + L.c;
      rn = (T)((double)a.r[e] + (-alpha) * (double)a.Pr[e]);
      a.r[e] = rn;
      if (fold) {
        a.x64[e] += (double)a.x[e];
        a.x[e] = (T)(alpha * (double)a.p[e]);
      } else {
        a.x[e] = (T)((double)a.x[e] + alpha * (double)a.p[e]);
      }
      part = (double)rn * (double)rn;
    }
    const unsigned m0 = a.l0mask[s];
    if (m0) {
      const T t0 = __shfl(rn, (int)L.r, kWave);
      const T t1 = __shfl(rn, (int)L.r + kSliceRows, kWave);
      const T t2 = __shfl(rn, (int)L.r + 2 * kSliceRows, kWave);
      if (i != ~0u && ((m0 >> L.r) & 1u)) {
        const T z = dia_row_dot(a.dinv0, L, t0, t1, t2);
        a.Pr[kVSB * (size_t)i + L.c] = z;
        if ((a.isomask[s] >> L.r) & 1u) srp = (double)rn * (double)z;
      }
    }
    part = wave_reduce_sum(part);
    if (L.lane == 0) a.partA2[s] = part;
    if (a.any_iso) {
      srp = wave_reduce_sum(srp);
      if (L.lane == 0) a.partI[s] = srp;
    }
  }
}

// forward level: Pr = D^-1 (r - L Pr); PART: partial (r, Pr) of the rows
// with no upper blocks (partF)
template <typename T, class V, bool PART>
__global__ void __launch_bounds__(V::WPB * kWave)
kb_fwd(BArgs<T> a, unsigned slice0, unsigned nsl) {
  const unsigned s0 = wave_first<V>(slice0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= slice0 + nsl) return;
    const SliceLane L = lane_of<V>(s);
    const RowMeta R = row_meta(a.F, L, s);
    if (ld_state(&a.M[R.mesh].done)) continue;
    const unsigned i = R.i;
    const bool valid = i != ~0u;
    T t = 0, rv = 0;
    Trip<T> d{};
    if (valid) {  // r_i and the D_i^-1 row first: off the gathers' chain
      rv = a.r[kVSB * (size_t)i + L.c];
      d = load_trip<T, sweep_nt<T>()>(a.f_dia, s, L.lane);
      t = sliced_row_sum_m<T, T, true, V::UW, kVSB, sweep_nt<T>()>(
          rv, (T)1, L, R.len, R.len, R.kmax, R.off, a.F.col, a.f_val, a.Pr);
    }
    const T t0 = __shfl(t, (int)L.r, kWave);
    const T t1 = __shfl(t, (int)L.r + kSliceRows, kWave);
    const T t2 = __shfl(t, (int)L.r + 2 * kSliceRows, kWave);
    double part = 0.0;
    if (valid) {
      const T out = (T)d.a * t0 + (T)d.b * t1 + (T)d.c * t2;  // dia_row_dot_m's products
      a.Pr[kVSB * (size_t)i + L.c] = out;
      if (PART && ((a.fumask[s] >> L.r) & 1u)) part = (double)rv * (double)out;
    }
    if (PART) {
      part = wave_reduce_sum(part);
      if (L.lane == 0) a.partF[s] = part;
    }
  }
}

// backward level: Pr -= U Pr, partial (r, Pr) of the rows it finalises
template <typename T, class V>
__global__ void __launch_bounds__(V::WPB * kWave)
kb_bwd(BArgs<T> a, unsigned slice0, unsigned nsl) {
  const unsigned s0 = wave_first<V>(slice0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= slice0 + nsl) return;
    const SliceLane L = lane_of<V>(s);
    const RowMeta R = row_meta(a.B, L, s);
    if (ld_state(&a.M[R.mesh].done)) continue;
    const unsigned i = R.i;
    double part = 0.0;
    if (i != ~0u) {
      const size_t e = kVSB * (size_t)i + L.c;
      const T rv = a.r[e];
      T t = a.Pr[e];
      t = sliced_row_sum_m<T, T, true, V::UW, kVSB, sweep_nt<T>()>(
          t, (T)1, L, R.len, R.len, R.kmax, R.off, a.B.col, a.b_val, a.Pr);
      a.Pr[e] = t;
      part = (double)rv * (double)t;
    }
    part = wave_reduce_sum(part);
    if (L.lane == 0) a.partB[s] = part;
  }
}

// p = beta p + Pr (mode 1) or p = Pr (mode 0)
template <typename T, class V>
__global__ void __launch_bounds__(V::WPB * kWave)
kb_pupdate(BArgs<T> a, unsigned nsl, int mode) {
  const unsigned s0 = wave_first<V>(0);
  for (int k = 0; k < V::SPW; ++k) {
    const unsigned s = s0 + k;
    if (s >= nsl) return;
    const MeshScal *S = a.M + a.A.mesh[s];
    const int dn = ld_state(&S->done);  // requested together with beta
    const double beta = ld_state(&S->beta);
    if (dn) continue;
    const SliceLane L = lane_of<V>(s);
    const unsigned i = L.c < 3 ? a.A.row[s * kSliceRows + L.r] : ~0u;
    if (i == ~0u) continue;
    const size_t e = kVSB * (size_t)i + L.c;
    a.p[e] = mode == 0 ? a.Pr[e] : (T)(beta * (double)a.p[e] + (double)a.Pr[e]);
  }
}

// the FP32 form's solution: x64 + x32 (a mesh that stops at a reliable
// update keeps its last x32 part unfolded)
__global__ void kb_mixed_out(double *__restrict__ x, const double *__restrict__ x64,
                             const float *__restrict__ x32, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = x64[i] + (double)x32[i];
}

__global__ void kb_fill(double *y, double v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = v;
}

// y += x
__global__ void kb_add(double *__restrict__ y, const double *__restrict__ x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

// ---------------------------------------------------------------------------
// per-mesh scalar steps: one wavefront per mesh sums that mesh's slice
// partials (lane-strided, then a fixed butterfly) and updates its state

enum BMode {
  kBInit = 0,         // |r0|^2 (FP64 form)
  kBAlpha = 1,        // alpha = rPr / pAp
  kBConvBeta = 2,     // FP64 |r|^2: stop test; then beta
  kBRPr0 = 3,         // first (r, Pr) (prologue: every backward level)
  kBInitMixed = 5,    // |b|^2
  kBAlphaMixed = 6,   // reliable update (FP64 |b - A x|^2, drift) when due; then alpha
  kBBetaMixed = 7     // FP32 |r|^2: decide the next reliable update; then beta
};

__device__ inline void mark_done(MeshScal *S, BatchGlobal *G) {
  S->done = 1;
  atomicAdd(&G->ndone, 1u);
}

// lane-strided sum of part[q] over [q0, q1), in increasing q, with the
// loads of U strides issued before their adds (each load round is a full
// memory latency: the partials were just written by other XCDs)
__device__ inline double lane_sum(const double *part, unsigned q0, unsigned q1, unsigned lane) {
  constexpr unsigned U = 4;
  double t = 0.0;
  for (unsigned q = q0 + lane; q < q1; q += U * kWave) {
    double v[U];
#pragma unroll
    for (unsigned u = 0; u < U; ++u) v[u] = q + u * kWave < q1 ? part[q + u * kWave] : 0.0;
#pragma unroll
    for (unsigned u = 0; u < U; ++u)
      if (q + u * kWave < q1) t += v[u];
  }
  return t;
}

template <int MODE>
__global__ void __launch_bounds__(kMeshPerBlock * kWave) kb_scalar(BArgs<float> a) {
  // (the argument block is precision-independent here; the caller passes a
  // float-typed view of the same control pointers)
  const unsigned m = blockIdx.x * kMeshPerBlock + threadIdx.x / kWave;
  if (m >= a.nmesh) return;
  const unsigned lane = threadIdx.x & (kWave - 1);
  MeshScal *S = a.M + m;
  constexpr bool kInit = MODE == kBInit || MODE == kBInitMixed;
  if (!kInit && ld_state(&S->done)) return;
  const bool upd = ld_state(&S->upd) != 0;
  const unsigned a0 = a.a_mptr[m], a1 = a.a_mptr[m + 1];
  double t = 0.0, t2 = 0.0, t3 = 0.0;
  if (MODE == kBRPr0) {
    for (unsigned l = 0; l < a.nlev_b; ++l) {
      const unsigned g = l * a.nmesh + m;
      t += lane_sum(a.partB, a.b_gptr[g], a.b_gptr[g + 1], lane);
    }
  } else if (MODE == kBConvBeta || MODE == kBBetaMixed) {
    t = lane_sum(a.partA2, a0, a1, lane);
    for (unsigned l = 1; l < a.nlev_b; ++l) {
      const unsigned g = l * a.nmesh + m;
      t2 += lane_sum(a.partB, a.b_gptr[g], a.b_gptr[g + 1], lane);
    }
    for (unsigned l = 1; l < a.nlev_f; ++l) {
      if (!a.flev_u0[l]) continue;
      const unsigned g = l * a.nmesh + m;
      t2 += lane_sum(a.partF, a.f_gptr[g], a.f_gptr[g + 1], lane);
    }
    if (a.any_iso) t2 += lane_sum(a.partI, a0, a1, lane);
  } else {
    t = lane_sum(a.partA, a0, a1, lane);
    if (MODE == kBAlphaMixed && upd) {
      t2 = lane_sum(a.partA2, a0, a1, lane);
      t3 = lane_sum(a.partE, a0, a1, lane);
    }
  }
  t = wave_reduce_sum(t);
  t2 = wave_reduce_sum(t2);
  t3 = wave_reduce_sum(t3);
  if (lane != 0) return;
  BatchGlobal *G = a.G;
  const bool bad = !(t == t) || t > 1.0e308;
  if (MODE == kBInit || MODE == kBInitMixed) {
    S->it = 0;
    S->nupd = 0;
    S->upd = 0;
    S->final_ = 0;
    S->it_upd = 0;
    S->restart = 0;
    S->converged = 0;
    S->nan = 0;
    S->done = 0;
    S->stalled = 0;
    S->it_best = 0;
    S->best2 = t;
    S->rmax2 = t;
    const double sb = MODE == kBInit && ld_state(&G->keep_scale) ? ld_state(&S->sq_b) : 0.0;
    if (sb > 0.0) {
      S->inv_sq0 = 1.0 / sb;
      S->res = sqrt(t / sb);
    } else {
      S->sq_b = t;
      S->inv_sq0 = 1.0 / t;
      S->res = 1.0;
    }
    if (t < 1.0e-30) {
      S->converged = 1;
      S->res = 0.0;
      mark_done(S, G);
    } else if (bad) {
      S->nan = 1;
      mark_done(S, G);
    }
  } else if (MODE == kBAlpha) {
    S->alpha = ld_state(&S->rPr) / t;
  } else if (MODE == kBRPr0) {
    S->rPr = t;
  } else if (MODE == kBConvBeta) {
    const int it = ld_state(&S->it) + 1;
    S->it = it;
    S->res = sqrt(t * ld_state(&S->inv_sq0));
    if (ld_state(&S->res) < G->tol) {
      S->converged = 1;
      mark_done(S, G);
    } else if (bad) {
      S->nan = 1;
      mark_done(S, G);
    } else if (it >= G->max_nitr) {
      mark_done(S, G);
    } else {
      S->beta = t2 / ld_state(&S->rPr);
      S->rPr = t2;
    }
  } else if (MODE == kBAlphaMixed) {
    if (upd) {
      const int k = ld_state(&S->nupd) + 1;
      S->nupd = k;
      S->rmax2 = t2;
      S->res = sqrt(t2 / ld_state(&S->sq_b));
      if (ld_state(&S->res) < G->tol) {
        S->converged = 1;
        mark_done(S, G);
      } else if (!(t2 == t2) || t2 > 1.0e308) {
        S->nan = 1;
        mark_done(S, G);
      } else if (ld_state(&S->final_) || k >= G->max_upd) {
        mark_done(S, G);
      } else if (t2 < 0.25 * ld_state(&S->best2)) {
        S->best2 = t2;
        S->it_best = ld_state(&S->it);
      } else if (ld_state(&S->it) - ld_state(&S->it_best) >= kStallIters) {
        // the FP32 iteration no longer reduces this mesh's FP64 residual:
        // stop it; the host finishes it with the FP64 form
        S->stalled = 1;
        mark_done(S, G);
      }
    }
    if (upd && (!(t > 0.0) || t3 > 1.0e-2 * t2)) {  // drifted: restart the direction
      S->alpha = 0.0;
      S->restart = 1;
    } else {
      S->alpha = ld_state(&S->rPr) / t;
    }
  } else {  // kBBetaMixed
    const int it = ld_state(&S->it) + 1;
    S->it = it;
    S->upd = 0;
    if (bad) {
      S->nan = 1;
      mark_done(S, G);
      return;
    }
    if (t < G->delta2 * ld_state(&S->rmax2) || sqrt(t / ld_state(&S->sq_b)) < G->tol || it >= G->max_nitr ||
        (G->period > 0 && it - ld_state(&S->it_upd) >= G->period)) {
      S->upd = 1;
      S->final_ = it >= G->max_nitr ? 1 : 0;
      S->it_upd = it;
    }
    S->beta = ld_state(&S->restart) ? 0.0 : t2 / ld_state(&S->rPr);
    S->restart = 0;
    S->rPr = t2;
  }
}

}  // namespace
}  // namespace dfm2h

using namespace dfm2h;

// ---------------------------------------------------------------------------
// host side

template <typename T>
struct BatchBufs {
  DBuf<T> r, x, Pr, p;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> sig;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    sig.clear();
  }
  ~BatchBufs() { drop_graph(); }
};

struct dfm2h_batch_s {
  unsigned nmesh = 0, nblk = 0;
  dfm2h_bsr A = nullptr;
  dfm2h_ilu P = nullptr;
  uint64_t a_serial = 0, p_serial = 0;  // pattern serials the plans were built from
  bool has_b = false;
  std::vector<unsigned> node_offset;
  PackedPlan aplan, fplan, bplan;
  DBuf<unsigned> a_mesh, f_mesh, b_mesh, a_mptr, b_gptr;
  unsigned nlev_f = 0, nlev_b = 0;
  PackedVals<double> a64, f64, b64v;
  PackedVals<float> a32, f32, b32;
  // fused level-0 data (see BArgs)
  DBuf<unsigned> l0mask, isomask, row_l0, fumask, f_gptr;
  DBuf<unsigned char> flev_u0;
  std::vector<unsigned char> flev_u0_host;
  bool any_iso = false;
  PackedVals<double> d64;
  PackedVals<float> d32;
  BatchBufs<double> s64;
  BatchBufs<float> s32;
  DBuf<double> x64, b64, partA, partA2, partB, partF, partI, partE;
  DBuf<MeshScal> M;
  DBuf<BatchGlobal> G;
  int variant[2] = {-1, -1};  // [fp32] chosen launch shape, -1 = autotune
  unsigned period = 25;       // FP32 form: forced reliable-update period
  // meshes the last FP32-form solve finished with the FP64 form, and the
  // full residual b - A x it was formed from (allocated on the first one)
  std::vector<int> fell_back;
  DBuf<double> fb_r;
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_poll[2] = {nullptr, nullptr}, ev_t0 = nullptr, ev_t1 = nullptr;
  BatchGlobal *host_g = nullptr;  // pinned, 2 entries
  ~dfm2h_batch_s() {
    s64.drop_graph();
    s32.drop_graph();
    if (host_g) (void)hipHostFree(host_g);
    for (hipEvent_t e : {ev_in, ev_poll[0], ev_poll[1], ev_t0, ev_t1})
      if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace dfm2h {
namespace {

// groups of rows -> per-slice mesh ids
std::vector<unsigned> slice_mesh(const PackedPlan &P, unsigned nmesh, bool level_major) {
  std::vector<unsigned> out(P.nslices);
  const size_t ng = P.group_ptr.size() - 1;
  for (size_t g = 0; g < ng; ++g)
    for (unsigned s = P.group_ptr[g]; s < P.group_ptr[g + 1]; ++s)
      out[s] = level_major ? (unsigned)(g % nmesh) : (unsigned)g;
  return out;
}

// rows of each level split by mesh: groups[l * nmesh + m]
std::vector<std::vector<unsigned>> level_mesh_groups(const std::vector<unsigned> &levptr,
                                                     const std::vector<unsigned> &rows,
                                                     const std::vector<unsigned> &node_offset) {
  const size_t nm = node_offset.size() - 1;
  const size_t nl = levptr.size() - 1;
  std::vector<std::vector<unsigned>> groups(nl * nm);
  for (size_t l = 0; l < nl; ++l) {
    size_t m = 0;
    for (unsigned k = levptr[l]; k < levptr[l + 1]; ++k) {
      const unsigned i = rows[k];  // ascending inside a level
      while (i >= node_offset[m + 1]) ++m;
      groups[l * nm + m].push_back(i);
    }
  }
  return groups;
}

template <typename T>
BArgs<T> make_bargs(dfm2h_batch B, BatchBufs<T> &S, const PackedVals<T> &av,
                    const PackedVals<T> &fv, const PackedVals<T> &bv, const PackedVals<T> &dv) {
  BArgs<T> a{};
  a.nmesh = B->nmesh;
  a.nlev_b = B->nlev_b;
  a.nlev_f = B->nlev_f;
  a.any_iso = B->any_iso ? 1 : 0;
  a.l0mask = B->l0mask.p;
  a.isomask = B->isomask.p;
  a.fumask = B->fumask.p;
  a.dinv0 = dv.dia.p;
  a.flev_u0 = B->flev_u0.p;
  a.partF = B->partF.p;
  a.partI = B->partI.p;
  a.partE = B->partE.p;
  a.f_gptr = B->f_gptr.p;
  a.A = BPlan{B->aplan.off.p, B->aplan.kmax.p, B->aplan.row.p, B->aplan.len.p, B->aplan.col.p,
              B->a_mesh.p};
  a.F = BPlan{B->fplan.off.p, B->fplan.kmax.p, B->fplan.row.p, B->fplan.len.p, B->fplan.col.p,
              B->f_mesh.p};
  a.B = BPlan{B->bplan.off.p, B->bplan.kmax.p, B->bplan.row.p, B->bplan.len.p, B->bplan.col.p,
              B->b_mesh.p};
  a.a_val = av.val.p;
  a.a_dia = av.dia.p;
  a.f_val = fv.val.p;
  a.f_dia = fv.dia.p;
  a.b_val = bv.val.p;
  a.a64_val = B->a64.val.p;
  a.a64_dia = B->a64.dia.p;
  a.r = S.r.p;
  a.x = S.x.p;
  a.Pr = S.Pr.p;
  a.p = S.p.p;
  a.x64 = B->x64.p;
  a.b64 = B->b64.p;
  a.partA = B->partA.p;
  a.partA2 = B->partA2.p;
  a.partB = B->partB.p;
  a.M = B->M.p;
  a.G = B->G.p;
  a.a_mptr = B->a_mptr.p;
  a.b_gptr = B->b_gptr.p;
  return a;
}

// the scalar kernels only touch control pointers: view any BArgs as BArgs<float>
template <typename T>
BArgs<float> ctl_view(const BArgs<T> &a) {
  BArgs<float> c{};
  c.nmesh = a.nmesh;
  c.nlev_b = a.nlev_b;
  c.nlev_f = a.nlev_f;
  c.any_iso = a.any_iso;
  c.flev_u0 = a.flev_u0;
  c.f_gptr = a.f_gptr;
  c.partA = a.partA;
  c.partA2 = a.partA2;
  c.partB = a.partB;
  c.partF = a.partF;
  c.partI = a.partI;
  c.partE = a.partE;
  c.M = a.M;
  c.G = a.G;
  c.a_mptr = a.a_mptr;
  c.b_gptr = a.b_gptr;
  return c;
}

template <int MODE, typename T>
void launch_scalar(const BArgs<T> &a, hipStream_t s) {
  hipLaunchKernelGGL(kb_scalar<MODE>, dim3(ceil_div(a.nmesh, kMeshPerBlock)),
                     dim3(kMeshPerBlock * kWave), 0, s, ctl_view(a));
}

// first_level = 0: every level (prologue: (r,Pr) partials of every
// backward level in partB); 1: the fused iteration (levels >= 1, forward
// levels with rows that have no upper blocks write partF)
template <typename T, class V>
void launch_sweeps(const BArgs<T> &a, dfm2h_batch B, hipStream_t s, unsigned first_level) {
  const unsigned nm = B->nmesh;
  for (unsigned l = first_level; l < B->nlev_f; ++l) {
    const unsigned s0 = B->fplan.group_ptr[(size_t)l * nm];
    const unsigned nsl = B->fplan.group_ptr[(size_t)(l + 1) * nm] - s0;
    if (!nsl) continue;
    if (first_level > 0 && B->flev_u0_host[l])
      hipLaunchKernelGGL((kb_fwd<T, V, true>), dim3(grid_of<V>(nsl)), dim3(V::WPB * kWave), 0,
                         s, a, s0, nsl);
    else
      hipLaunchKernelGGL((kb_fwd<T, V, false>), dim3(grid_of<V>(nsl)), dim3(V::WPB * kWave), 0,
                         s, a, s0, nsl);
  }
  for (unsigned l = first_level; l < B->nlev_b; ++l) {
    const unsigned s0 = B->bplan.group_ptr[(size_t)l * nm];
    const unsigned nsl = B->bplan.group_ptr[(size_t)(l + 1) * nm] - s0;
    if (!nsl) continue;
    hipLaunchKernelGGL((kb_bwd<T, V>), dim3(grid_of<V>(nsl)), dim3(V::WPB * kWave), 0, s, a, s0,
                       nsl);
  }
}

template <typename T, class V>
void launch_prologue(const BArgs<T> &a, dfm2h_batch B, hipStream_t s) {
  const unsigned na = B->aplan.nslices;
  const dim3 ga(grid_of<V>(na)), bs(V::WPB * kWave);
  hipLaunchKernelGGL((kb_init<T, V>), ga, bs, 0, s, a, na);
  if (sizeof(T) == 4)
    launch_scalar<kBInitMixed>(a, s);
  else
    launch_scalar<kBInit>(a, s);
  launch_sweeps<T, V>(a, B, s, 0);
  launch_scalar<kBRPr0>(a, s);
  hipLaunchKernelGGL((kb_pupdate<T, V>), ga, bs, 0, s, a, na, 0);
}

// spmv (+ reliable update) -> alpha -> update (+ forward level 0) -> forward
// levels >= 1 -> backward levels >= 1 -> stop test / update decision + beta
// -> p update
template <typename T, class V>
void launch_iteration(const BArgs<T> &a, dfm2h_batch B, hipStream_t s) {
  const unsigned na = B->aplan.nslices;
  const dim3 ga(grid_of<V>(na)), bs(V::WPB * kWave);
  hipLaunchKernelGGL((kb_spmv<T, V>), ga, bs, 0, s, a, na);
  if (sizeof(T) == 4)
    launch_scalar<kBAlphaMixed>(a, s);
  else
    launch_scalar<kBAlpha>(a, s);
  hipLaunchKernelGGL((kb_update<T, V>), ga, bs, 0, s, a, na);
  launch_sweeps<T, V>(a, B, s, 1);
  if (sizeof(T) == 4)
    launch_scalar<kBBetaMixed>(a, s);
  else
    launch_scalar<kBConvBeta>(a, s);
  hipLaunchKernelGGL((kb_pupdate<T, V>), ga, bs, 0, s, a, na, 1);
}

template <typename T>
using LaunchFn = void (*)(const BArgs<T> &, dfm2h_batch, hipStream_t);

template <typename T>
LaunchFn<T> prologue_fn(int v) {
  switch (v) {
    case 0: return launch_prologue<T, V0>;
    case 1: return launch_prologue<T, V1>;
    case 2: return launch_prologue<T, V2>;
    case 3: return launch_prologue<T, V3>;
    case 4: return launch_prologue<T, V4>;
    default: return launch_prologue<T, V5>;
  }
}

template <typename T>
LaunchFn<T> iteration_fn(int v) {
  switch (v) {
    case 0: return launch_iteration<T, V0>;
    case 1: return launch_iteration<T, V1>;
    case 2: return launch_iteration<T, V2>;
    case 3: return launch_iteration<T, V3>;
    case 4: return launch_iteration<T, V4>;
    default: return launch_iteration<T, V5>;
  }
}

int batch_bucket(dfm2h_batch B) { return dof_bucket(3 * (size_t)B->nblk); }

// values packed for the batch plans, from the current A and P
template <typename T>
void ensure_batch_vals(dfm2h_batch B, PackedVals<T> &av, PackedVals<T> &fv, PackedVals<T> &bv,
                       PackedVals<T> &dv) {
  dfm2h_bsr A = B->A;
  dfm2h_ilu P = B->P;
  if (sizeof(T) == 4)  // the FP32 copy in the difference form (kb_spmv)
    ensure_packed_rowsum(B->aplan, av, A->vcrs.p, A->vdia.p, A->pat.colind_d.p, A->version,
                         B->s);
  else
    ensure_packed(B->aplan, av, A->vcrs.p, A->vdia.p, A->version, B->s);
  ensure_packed(B->fplan, fv, P->vcrs.p, P->vdia.p, P->version, B->s);
  ensure_packed(B->bplan, bv, P->vcrs.p, nullptr, P->version, B->s);
  // D^-1 of the forward-level-0 rows in the matrix-slice layout
  if (dv.version != P->version || !dv.dia.p) {
    const unsigned nsl = B->aplan.nslices;
    dv.dia.alloc(std::max<size_t>((size_t)nsl * kSliceVals, 1));
    if (nsl) {
      hipLaunchKernelGGL(k_pack_dia<T>, dim3(pack_grid(nsl)), dim3(256), 0, B->s, nsl,
                         B->row_l0.p, P->vdia.p, dv.dia.p);
      check_launch("k_pack_dia");
    }
    dv.version = P->version;
  }
}

template <typename T>
void ensure_batch_bufs(dfm2h_batch B, BatchBufs<T> &S) {
  const size_t n = (size_t)kVSB * B->nblk;
  if (S.r.n != n) {
    S.r.alloc(n);
    S.x.alloc(n);
    S.Pr.alloc(n);
    S.p.alloc(n);
  }
}

// Runs one whole solve of every mesh with variant v on the batch stream;
// b is already in B->b64.  Returns when every mesh is done.
template <typename T>
void run_batch(dfm2h_batch B, BatchBufs<T> &S, const BArgs<T> &a, int v, double tol,
               double delta, unsigned max_nitr, int keep_scale = 0) {
  hipStream_t s = B->s;
  BatchGlobal g{};
  g.keep_scale = keep_scale;
  g.tol = tol;
  g.delta2 = delta * delta;
  g.max_nitr = (int)max_nitr;
  g.max_upd = (int)max_nitr;
  g.ndone = 0;
  g.period = (int)B->period;
  DFM2H_HIP(hipMemcpyAsync(B->G.p, &g, sizeof(BatchGlobal), hipMemcpyHostToDevice, s));
  prologue_fn<T>(v)(a, B, s);
  check_launch("batch prologue");
  const unsigned chunk = std::max(1u, std::min(kChunk, max_nitr));
  std::vector<uintptr_t> sig = {(uintptr_t)v,        (uintptr_t)chunk,    (uintptr_t)a.a_val,
                                (uintptr_t)a.a_dia,  (uintptr_t)a.f_val,  (uintptr_t)a.f_dia,
                                (uintptr_t)a.b_val,  (uintptr_t)a.a64_val, (uintptr_t)a.a64_dia,
                                (uintptr_t)a.r,      (uintptr_t)a.x64,    (uintptr_t)a.b64,
                                (uintptr_t)a.partA,  (uintptr_t)a.M,       (uintptr_t)a.dinv0};
  if (!S.exec || S.sig != sig) {
    S.drop_graph();
    DFM2H_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (unsigned k = 0; k < chunk; ++k) iteration_fn<T>(v)(a, B, s);
    hipError_t e = hipStreamEndCapture(s, &S.graph);
    DFM2H_REQUIRE(e == hipSuccess, DFM2H_ERR_HIP,
                  std::string("graph capture: ") + hipGetErrorString(e));
    DFM2H_HIP(hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
    S.sig = sig;
  }
  // the FP32 form may take one extra chunk for the final reliable update
  const unsigned nchunks = (max_nitr + chunk - 1) / chunk + (sizeof(T) == 4 ? 1 : 0);
  bool finished = false;
  for (unsigned c = 0; c < nchunks && !finished; ++c) {
    DFM2H_HIP(hipGraphLaunch(S.exec, s));
    DFM2H_HIP(hipMemcpyAsync(&B->host_g[c & 1], B->G.p, sizeof(BatchGlobal),
                             hipMemcpyDeviceToHost, s));
    DFM2H_HIP(hipEventRecord(B->ev_poll[c & 1], s));
    if (c > 0) {
      DFM2H_HIP(hipEventSynchronize(B->ev_poll[(c - 1) & 1]));
      finished = B->host_g[(c - 1) & 1].ndone >= B->nmesh;
    }
  }
  DFM2H_HIP(hipStreamSynchronize(s));
}

template <typename T>
BArgs<T> prepare(dfm2h_batch B, BatchBufs<T> &S);

template <>
BArgs<double> prepare<double>(dfm2h_batch B, BatchBufs<double> &S) {
  ensure_batch_bufs(B, S);
  ensure_batch_vals(B, B->a64, B->f64, B->b64v, B->d64);
  return make_bargs<double>(B, S, B->a64, B->f64, B->b64v, B->d64);
}

template <>
BArgs<float> prepare<float>(dfm2h_batch B, BatchBufs<float> &S) {
  ensure_batch_bufs(B, S);
  ensure_batch_vals(B, B->a32, B->f32, B->b32, B->d32);
  ensure_packed(B->aplan, B->a64, B->A->vcrs.p, B->A->vdia.p, B->A->version, B->s);
  return make_bargs<float>(B, S, B->a32, B->f32, B->b32, B->d32);
}

// Times a short fixed-length solve (tolerance 0, `nit` iterations per
// mesh) for every variant and returns the fastest; ms[v] = time per
// iteration of variant v.
template <typename T>
int autotune(dfm2h_batch B, BatchBufs<T> &S, double *ms, size_t ms_cap) {
  const unsigned nit = 2 * kChunk;
  int best = 0;
  double best_ms = 1e300;
  for (int v = 0; v < kNumVariants; ++v) {
    const BArgs<T> a = prepare<T>(B, S);
    double t_ms = 0.0;
    for (int rep = 0; rep < 2; ++rep) {  // first run captures and warms up
      DFM2H_HIP(hipEventRecord(B->ev_t0, B->s));
      run_batch<T>(B, S, a, v, 0.0, 1e-2, nit);
      DFM2H_HIP(hipEventRecord(B->ev_t1, B->s));
      DFM2H_HIP(hipEventSynchronize(B->ev_t1));
      float t = 0.f;
      DFM2H_HIP(hipEventElapsedTime(&t, B->ev_t0, B->ev_t1));
      t_ms = (double)t / nit;
    }
    if (ms && (size_t)v < ms_cap) ms[v] = t_ms;
    if (t_ms < best_ms) {
      best_ms = t_ms;
      best = v;
    }
  }
  S.drop_graph();
  tune_store("batch", batch_bucket(B), sizeof(T) == 4 ? 1 : 0, best, best_ms);
  return best;
}

void need_batch(dfm2h_batch B) {
  DFM2H_REQUIRE(B, DFM2H_ERR_INVALID, "null handle");
  DFM2H_REQUIRE(B->A->pattern_serial == B->a_serial && B->P->pattern_serial == B->p_serial,
                DFM2H_ERR_STATE, "matrix / preconditioner pattern changed since batch_create");
  DFM2H_REQUIRE(B->P->decomposed, DFM2H_ERR_STATE, "preconditioner not decomposed");
  flush_merge(B->A);
}

}  // namespace
}  // namespace dfm2h

extern "C" {

int dfm2h_batch_create(dfm2h_batch *out, dfm2h_bsr A, dfm2h_ilu P, const unsigned *node_offset,
                       size_t nmesh) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out && A && P && node_offset, DFM2H_ERR_INVALID, "null argument");
    *out = nullptr;
    DFM2H_REQUIRE(A->has_pattern && P->has_pattern, DFM2H_ERR_STATE, "pattern not set");
    DFM2H_REQUIRE(A->nblk == P->nblk, DFM2H_ERR_INVALID, "matrix / preconditioner size mismatch");
    DFM2H_REQUIRE(nmesh >= 1 && nmesh < (1u << 24), DFM2H_ERR_INVALID, "bad mesh count");
    const unsigned nblk = A->nblk;
    std::vector<unsigned> off(node_offset, node_offset + nmesh + 1);
    DFM2H_REQUIRE(off[0] == 0 && off[nmesh] == nblk, DFM2H_ERR_INVALID,
                  "node_offset must start at 0 and end at the node count");
    for (size_t m = 0; m < nmesh; ++m)
      DFM2H_REQUIRE(off[m] <= off[m + 1], DFM2H_ERR_INVALID, "node_offset must be non-decreasing");
    // every block of row i must stay inside i's mesh (the union is block diagonal)
    {
      size_t m = 0;
      for (unsigned i = 0; i < nblk; ++i) {
        while (i >= off[m + 1]) ++m;
        for (unsigned q = A->pat.colind[i]; q < A->pat.colind[i + 1]; ++q) {
          const unsigned j = A->pat.rowptr[q];
          DFM2H_REQUIRE(j >= off[m] && j < off[m + 1], DFM2H_ERR_PATTERN,
                        "matrix couples two meshes (row " + std::to_string(i) + ", column " +
                            std::to_string(j) + ")");
        }
      }
    }
    auto B = std::make_unique<dfm2h_batch_s>();
    B->nmesh = (unsigned)nmesh;
    B->nblk = nblk;
    B->A = A;
    B->P = P;
    B->a_serial = A->pattern_serial;
    B->p_serial = P->pattern_serial;
    B->node_offset = off;
    // matrix plan: one group per mesh
    {
      std::vector<std::vector<unsigned>> groups(nmesh);
      for (size_t m = 0; m < nmesh; ++m)
        for (unsigned i = off[m]; i < off[m + 1]; ++i) groups[m].push_back(i);
      std::vector<unsigned> s0(nblk), s1(nblk);
      for (unsigned i = 0; i < nblk; ++i) {
        s0[i] = A->pat.colind[i];
        s1[i] = A->pat.colind[i + 1];
      }
      B->aplan.build(groups, s0, s1, A->pat.rowptr);
      B->a_mesh.upload(slice_mesh(B->aplan, B->nmesh, false));
      B->a_mptr.upload(B->aplan.group_ptr);
      // forward-level-0 rows (no lower blocks) and isolated rows per matrix slice
      std::vector<unsigned> l0(B->aplan.nslices, 0), iso(B->aplan.nslices, 0);
      std::vector<unsigned> rl0(std::max<size_t>((size_t)B->aplan.nslices * kSliceRows, 1), ~0u);
      for (size_t m = 0; m < nmesh; ++m)
        for (unsigned q = B->aplan.group_ptr[m]; q < B->aplan.group_ptr[m + 1]; ++q)
          for (unsigned r = 0; r < (unsigned)kSliceRows; ++r) {
            const size_t i = off[m] + (size_t)(q - B->aplan.group_ptr[m]) * kSliceRows + r;
            if (i >= off[m + 1] || P->diaind[i] != P->colind[i]) continue;
            l0[q] |= 1u << r;
            rl0[(size_t)q * kSliceRows + r] = (unsigned)i;
            if (P->diaind[i] == P->colind[i + 1]) iso[q] |= 1u << r;
          }
      B->any_iso = std::any_of(iso.begin(), iso.end(), [](unsigned v) { return v != 0; });
      B->l0mask.upload(l0);
      B->isomask.upload(iso);
      B->row_l0.upload(rl0);
    }
    // sweep plans: one group per (level, mesh), level-major
    {
      std::vector<unsigned> levptr, rows, s0(nblk), s1(nblk);
      build_levels(nblk, P->colind, P->rowptr, P->diaind, true, levptr, rows);
      B->nlev_f = (unsigned)levptr.size() - 1;
      for (unsigned i = 0; i < nblk; ++i) {
        s0[i] = P->colind[i];
        s1[i] = P->diaind[i];
      }
      const auto fgroups = level_mesh_groups(levptr, rows, off);
      B->fplan.build(fgroups, s0, s1, P->rowptr);
      B->f_mesh.upload(slice_mesh(B->fplan, B->nmesh, true));
      B->f_gptr.upload(B->fplan.group_ptr);
      // per forward slice: rows with no upper blocks; per level: any such row
      std::vector<unsigned> um;
      B->flev_u0_host.assign(B->nlev_f, 0);
      for (size_t g = 0; g < fgroups.size(); ++g)
        for (size_t b = 0; b < fgroups[g].size(); b += kSliceRows) {
          unsigned msk = 0;
          for (size_t k = b; k < std::min(fgroups[g].size(), b + kSliceRows); ++k) {
            const unsigned i = fgroups[g][k];
            if (P->diaind[i] == P->colind[i + 1]) msk |= 1u << (k - b);
          }
          um.push_back(msk);
          if (msk) B->flev_u0_host[g / nmesh] = 1;
        }
      B->fumask.upload(um);
      B->flev_u0.upload(B->flev_u0_host);
      build_levels(nblk, P->colind, P->rowptr, P->diaind, false, levptr, rows);
      B->nlev_b = (unsigned)levptr.size() - 1;
      for (unsigned i = 0; i < nblk; ++i) {
        s0[i] = P->diaind[i];
        s1[i] = P->colind[i + 1];
      }
      B->bplan.build(level_mesh_groups(levptr, rows, off), s0, s1, P->rowptr);
      B->b_mesh.upload(slice_mesh(B->bplan, B->nmesh, true));
      B->b_gptr.upload(B->bplan.group_ptr);
    }
    const size_t n = (size_t)kVSB * nblk;
    B->x64.alloc(std::max<size_t>(n, 1));
    B->b64.alloc(std::max<size_t>(n, 1));
    B->partA.alloc(std::max<size_t>(B->aplan.nslices, 1));
    B->partA2.alloc(std::max<size_t>(B->aplan.nslices, 1));
    B->partB.alloc(std::max<size_t>(B->bplan.nslices, 1));
    B->partF.alloc(std::max<size_t>(B->fplan.nslices, 1));
    B->partI.alloc(std::max<size_t>(B->aplan.nslices, 1));
    B->partE.alloc(std::max<size_t>(B->aplan.nslices, 1));
    B->M.alloc(nmesh);
    B->G.alloc(1);
    DFM2H_HIP(hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking));
    DFM2H_HIP(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&B->ev_poll[0], hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&B->ev_poll[1], hipEventDisableTiming));
    DFM2H_HIP(hipEventCreate(&B->ev_t0));
    DFM2H_HIP(hipEventCreate(&B->ev_t1));
    DFM2H_HIP(hipHostMalloc((void **)&B->host_g, 2 * sizeof(BatchGlobal), hipHostMallocDefault));
    DFM2H_HIP(hipStreamSynchronize(stream()));
    *out = B.release();
  });
}

int dfm2h_batch_destroy(dfm2h_batch B) {
  return guard(__func__, [&] { delete B; });
}

int dfm2h_batch_info(dfm2h_batch B, size_t *nmesh, size_t *nslices, unsigned *n_forward,
                     unsigned *n_backward) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(B, DFM2H_ERR_INVALID, "null handle");
    if (nmesh) *nmesh = B->nmesh;
    if (nslices) *nslices = B->aplan.nslices;
    if (n_forward) *n_forward = B->nlev_f;
    if (n_backward) *n_backward = B->nlev_b;
  });
}

int dfm2h_batch_set_update_period(dfm2h_batch B, unsigned update_period) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(B, DFM2H_ERR_INVALID, "null handle");
    B->period = update_period;
  });
}

int dfm2h_batch_set_variant(dfm2h_batch B, int fp32, int variant) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(B, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(variant >= -1 && variant < kNumVariants, DFM2H_ERR_INVALID, "bad variant");
    B->variant[fp32 ? 1 : 0] = variant;
  });
}

int dfm2h_batch_get_variant(dfm2h_batch B, int fp32, int *variant, const char **name) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(B && variant, DFM2H_ERR_INVALID, "null argument");
    const int v = B->variant[fp32 ? 1 : 0];
    *variant = v;
    if (name) *name = v >= 0 ? kVariantNames[v] : "autotune";
  });
}

int dfm2h_batch_tune(dfm2h_batch B, int fp32, int *best, double *ms, size_t ms_cap,
                     size_t *n_variants) {
  return guard(__func__, [&] {
    need_batch(B);
    DFM2H_HIP(hipEventRecord(B->ev_in, stream()));
    DFM2H_HIP(hipStreamWaitEvent(B->s, B->ev_in, 0));
    const size_t n = (size_t)kVSB * B->nblk;
    if (!B->has_b && n) {  // no right-hand side seen yet: time on b = 1
      hipLaunchKernelGGL(kb_fill, dim3(std::min<size_t>(2048, ceil_div(n, 256))), dim3(256), 0,
                         B->s, B->b64.p, 1.0, n);
      check_launch("kb_fill");
    }
    const int v = fp32 ? autotune<float>(B, B->s32, ms, ms_cap)
                       : autotune<double>(B, B->s64, ms, ms_cap);
    B->variant[fp32 ? 1 : 0] = v;
    if (best) *best = v;
    if (n_variants) *n_variants = kNumVariants;
  });
}

int dfm2h_batch_fell_back(dfm2h_batch B, int *flags) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(B && flags, DFM2H_ERR_INVALID, "null argument");
    for (unsigned m = 0; m < B->nmesh; ++m)
      flags[m] = m < B->fell_back.size() ? B->fell_back[m] : 0;
  });
}

int dfm2h_batch_solve(dfm2h_batch B, const double *b, double *x, double tol, int fp32,
                      double delta, unsigned max_nitr, unsigned *iters, double *relres,
                      int *state) {
  return guard(__func__, [&] {
    need_batch(B);
    DFM2H_REQUIRE(b && x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(b);
    DFM2H_DEVICE_ARG(x);
    DFM2H_REQUIRE(max_nitr >= 1, DFM2H_ERR_INVALID, "max_nitr must be >= 1");
    DFM2H_REQUIRE(!fp32 || (delta > 0.0 && delta < 1.0), DFM2H_ERR_INVALID,
                  "delta must be in (0, 1)");
    const size_t n = (size_t)kVSB * B->nblk;
    DFM2H_HIP(hipEventRecord(B->ev_in, stream()));
    DFM2H_HIP(hipStreamWaitEvent(B->s, B->ev_in, 0));
    if (n)
      DFM2H_HIP(hipMemcpyAsync(B->b64.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, B->s));
    B->has_b = true;
    const int p = fp32 ? 1 : 0;
    int v = B->variant[p];
    if (v < 0) v = tune_lookup("batch", batch_bucket(B), p, kNumVariants);
    if (fp32) {
      if (v < 0) v = autotune<float>(B, B->s32, nullptr, 0);
      const BArgs<float> a = prepare<float>(B, B->s32);
      run_batch<float>(B, B->s32, a, v, tol, delta, max_nitr);
      if (n) {
        hipLaunchKernelGGL(kb_mixed_out, dim3(std::min<size_t>(2048, ceil_div(n, 256))),
                           dim3(256), 0, B->s, x, B->x64.p, B->s32.x.p, n);
        check_launch("kb_mixed_out");
      }
    } else {
      if (v < 0) v = autotune<double>(B, B->s64, nullptr, 0);
      const BArgs<double> a = prepare<double>(B, B->s64);
      run_batch<double>(B, B->s64, a, v, tol, 0.5, max_nitr);
      if (n)
        DFM2H_HIP(hipMemcpyAsync(x, B->s64.x.p, n * sizeof(double), hipMemcpyDeviceToDevice,
                                 B->s));
    }
    B->variant[p] = v;
    std::vector<MeshScal> h(B->nmesh);
    DFM2H_HIP(hipMemcpyAsync(h.data(), B->M.p, h.size() * sizeof(MeshScal),
                             hipMemcpyDeviceToHost, B->s));
    DFM2H_HIP(hipStreamSynchronize(B->s));
    B->fell_back.assign(B->nmesh, 0);
    std::vector<unsigned> stalled;
    for (unsigned m = 0; m < B->nmesh; ++m)
      if (fp32 && h[m].stalled && !h[m].nan) stalled.push_back(m);
    if (!stalled.empty()) {
      // The FP32 iteration stopped reducing these meshes' FP64 residuals:
      // finish them with the FP64 form on the correction, e = A^-1 (b - A x)
      // with the right-hand side zero on every other mesh (they stop at
      // once) and each stop test relative to the mesh's own |b| (kept in
      // sq_b); then x += e.
      int v64 = B->variant[0];
      if (v64 < 0) v64 = tune_lookup("batch", batch_bucket(B), 0, kNumVariants);
      if (v64 < 0) v64 = autotune<double>(B, B->s64, nullptr, 0);  // before b64 is set
      B->variant[0] = v64;
      if (B->fb_r.n != n) B->fb_r.alloc(n);
      DFM2H_HIP(hipMemcpyAsync(B->fb_r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice,
                               stream()));
      const int rc = dfm2h_bsr_matvec(B->A, B->fb_r.p, -1.0, x, 1.0);
      DFM2H_REQUIRE(rc == DFM2H_OK, rc, dfm2h_last_error());
      DFM2H_HIP(hipEventRecord(B->ev_in, stream()));
      DFM2H_HIP(hipStreamWaitEvent(B->s, B->ev_in, 0));
      hipLaunchKernelGGL(kb_fill, dim3(std::min<size_t>(2048, ceil_div(n, 256))), dim3(256), 0,
                         B->s, B->b64.p, 0.0, n);
      check_launch("kb_fill");
      for (unsigned m : stalled) {
        const size_t r0 = (size_t)kVSB * B->node_offset[m];
        const size_t r1 = (size_t)kVSB * B->node_offset[m + 1];
        DFM2H_HIP(hipMemcpyAsync(B->b64.p + r0, B->fb_r.p + r0, (r1 - r0) * sizeof(double),
                                 hipMemcpyDeviceToDevice, B->s));
        B->fell_back[m] = 1;
      }
      const BArgs<double> a = prepare<double>(B, B->s64);
      run_batch<double>(B, B->s64, a, v64, tol, 0.5, max_nitr, 1);
      hipLaunchKernelGGL(kb_add, dim3(std::min<size_t>(2048, ceil_div(n, 256))), dim3(256), 0,
                         B->s, x, (const double *)B->s64.x.p, n);
      check_launch("kb_add");
      std::vector<MeshScal> hf(B->nmesh);
      DFM2H_HIP(hipMemcpyAsync(hf.data(), B->M.p, hf.size() * sizeof(MeshScal),
                               hipMemcpyDeviceToHost, B->s));
      DFM2H_HIP(hipStreamSynchronize(B->s));
      for (unsigned m : stalled) {
        hf[m].it += h[m].it;  // FP32 iterations, then the FP64 form's
        h[m] = hf[m];
      }
    }
    bool any_nan = false;
    for (unsigned m = 0; m < B->nmesh; ++m) {
      if (iters) iters[m] = (unsigned)h[m].it;
      if (relres) relres[m] = h[m].res;
      if (state) state[m] = h[m].nan ? -1 : h[m].converged ? 1 : 0;
      any_nan |= h[m].nan != 0;
    }
    DFM2H_REQUIRE(!any_nan, DFM2H_ERR_NAN, "a mesh's residual became NaN or Inf");
  });
}

}  // extern "C"
