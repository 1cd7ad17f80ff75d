// pcg.hip -- view_vectorx.h operator set and the fused Solve_PCG
// (lsitrsol.h:80-148) in FP64, plus FP32-inner / FP64-refinement.
//
// Fused iteration (one HIP graph node per kernel, K iterations per graph):
//   spmv   Ap = A p, one partial (p,Ap) per workgroup
//   scal   alpha = rPr/pAp                          (one workgroup)
//   update r += -alpha Ap, x += alpha p, partial (r,r)
//   scal   history, convergence test, iteration count
//   fwd    one kernel per dependency level: Pr = L^-1 r (reads r directly,
//          which fuses the reference's copy Pr = r)
//   bwd    one kernel per level: Pr = U^-1 Pr, partial (r,Pr)
//   scal   beta = rPr1/rPr, rPr = rPr1
//   pupd   p = beta p + Pr
// Every kernel returns at once when the device-side `done` flag is set, so
// iterations launched after convergence cost only their launch.  Scalars
// never leave HBM; the host polls `done` once per chunk, one chunk behind.
// All grid reductions use a fixed partition and a fixed final order, hence
// are deterministic run to run.
#include "handles.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <mutex>
#include <type_traits>

namespace dfm2h {

struct PcgScal {
  double rPr;  // (r, M^-1 r)
  double alpha;
  double beta;
  double inv_sq0;
  double tol;
  double sq;
  int it;
  int done;
  int converged;
  int max_nitr;
  int nan;
  int pad;
  // mixed precision (reliable updates)
  double sq_b;    // |b|^2 in FP64
  double rmax2;   // FP64 |r|^2 at the last reliable update
  double delta2;  // update when the FP32 |r|^2 fell below delta2 * rmax2
  int upd;        // the next SpMV recomputes r in FP64, the next update folds x32
  int nupd;       // FP64 residuals recorded after the initial |b|
  int max_upd;
  int final_;     // iteration cap reached: stop after this update
  int it_upd;     // inner iteration of the last reliable update
  int period;     // force an update after this many inner iterations (0: off)
  int restart;    // the search direction restarts from the true residual
  int pad3;
  double best2;   // smallest FP64 |r|^2 of a reliable update that at least halved |r|
  int it_best;    // inner iteration of that update
  int stalled;    // kStallIters inner iterations since it: the FP32 iteration stopped converging
};

// Field by field copy of the state (a whole-struct copy keeps a per-thread
// stack copy alive, which the compiler then places in LDS for every thread).
__device__ __attribute__((always_inline)) inline void copy_state(PcgScal &d, const PcgScal &s) {
  d.rPr = s.rPr;
  d.alpha = s.alpha;
  d.beta = s.beta;
  d.inv_sq0 = s.inv_sq0;
  d.tol = s.tol;
  d.sq = s.sq;
  d.it = s.it;
  d.done = s.done;
  d.converged = s.converged;
  d.max_nitr = s.max_nitr;
  d.nan = s.nan;
  d.pad = s.pad;
  d.sq_b = s.sq_b;
  d.rmax2 = s.rmax2;
  d.delta2 = s.delta2;
  d.upd = s.upd;
  d.nupd = s.nupd;
  d.max_upd = s.max_upd;
  d.final_ = s.final_;
  d.it_upd = s.it_upd;
  d.period = s.period;
  d.restart = s.restart;
  d.pad3 = s.pad3;
  d.best2 = s.best2;
  d.it_best = s.it_best;
  d.stalled = s.stalled;
}

__device__ __attribute__((always_inline)) inline void load_state(PcgScal &d, const PcgScal *s) {
  d.rPr = ld_state(&s->rPr);
  d.alpha = ld_state(&s->alpha);
  d.beta = ld_state(&s->beta);
  d.inv_sq0 = ld_state(&s->inv_sq0);
  d.tol = ld_state(&s->tol);
  d.sq = ld_state(&s->sq);
  d.it = ld_state(&s->it);
  d.done = ld_state(&s->done);
  d.converged = ld_state(&s->converged);
  d.max_nitr = ld_state(&s->max_nitr);
  d.nan = ld_state(&s->nan);
  d.pad = ld_state(&s->pad);
  d.sq_b = ld_state(&s->sq_b);
  d.rmax2 = ld_state(&s->rmax2);
  d.delta2 = ld_state(&s->delta2);
  d.upd = ld_state(&s->upd);
  d.nupd = ld_state(&s->nupd);
  d.max_upd = ld_state(&s->max_upd);
  d.final_ = ld_state(&s->final_);
  d.it_upd = ld_state(&s->it_upd);
  d.period = ld_state(&s->period);
  d.restart = ld_state(&s->restart);
  d.pad3 = ld_state(&s->pad3);
  d.best2 = ld_state(&s->best2);
  d.it_best = ld_state(&s->it_best);
  d.stalled = ld_state(&s->stalled);
}

constexpr int kVecBlock = 256;
constexpr unsigned kVecGridMax = 2048;
constexpr int kScalBlock = 1024;  // the one-workgroup scalar steps
// node stride of the fused solvers' internal vectors: 3, the caller's
// layout (measured faster on MI355X than a padded stride of 4)
constexpr int kVS = 3;

static unsigned vec_grid(size_t n) {
  return std::max(1u, std::min(kVecGridMax, ceil_div(n, kVecBlock)));
}


// ---------------------------------------------------------------------------
// generic deterministic dot (host-facing Dot)

__global__ void __launch_bounds__(kVecBlock)
k_dot(const double *__restrict__ x, const double *__restrict__ y, size_t n, double *partials) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kVecBlock)
    s += x[i] * y[i];
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kVecBlock)
k_sum_partials(const double *__restrict__ partials, unsigned n, double *out) {
  const double t = reduce_partials<kVecBlock>(partials, n);
  if (threadIdx.x == 0) *out = t;
}

__global__ void k_add_scaled(double *y, double alpha, const double *x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

__global__ void k_scale_and_add(double *y, double beta, const double *x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = beta * y[i] + x[i];
}

template <typename T>
__global__ void k_fill(T *y, T v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = v;
}

template <typename TO, typename TI>
__global__ void k_convert(TO *__restrict__ y, const TI *__restrict__ x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}

// node stride 3 (caller's vectors) <-> kVS (the fused solvers' vectors)
template <typename TO, typename TI>
__global__ void k_pad4(TO *__restrict__ y4, const TI *__restrict__ x3, size_t nnode) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnode * kVS;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t node = i / kVS;
    const unsigned c = (unsigned)(i % kVS);
    y4[i] = c < 3 ? (TO)x3[node * 3 + c] : (TO)0;
  }
}

template <typename TO, typename TI>
__global__ void k_unpad4(TO *__restrict__ y3, const TI *__restrict__ x4, size_t nnode) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnode * 3;
       i += (size_t)gridDim.x * blockDim.x)
    y3[i] = (TO)x4[(i / 3) * kVS + i % 3];
}

// ---------------------------------------------------------------------------
// fused PCG kernels
//
// Two launches of the reference's sweep structure are folded into
// neighbours, without changing any arithmetic:
//  * forward level 0 (rows with no lower blocks: Pr_i = D_i^-1 r_i) is
//    computed by the kernel that produces r (the r/x update, or the FP64
//    residual of a reliable update), from the row's three r components
//    exchanged inside the wavefront;
//  * backward level 0 (rows with no upper blocks) changes nothing; its share
//    of (r, Pr) is summed where those rows' Pr is final (the forward level
//    that computes them).
// Every partial of (r, Pr) of an iteration goes to `partials3`:
//   [backward levels >= 1 | forward levels with final rows | update].

// T: precision of the Krylov vectors; M: storage precision of the matrix
// and factor values the SpMV and sweeps stream (widened to T before use).
// M = FP32 is the mixed solver (FP64 reliable updates); with T = FP32 its
// solution is kept split as x64 + x32.
template <typename T, typename M>
constexpr bool is_mixed() {
  return sizeof(M) == 4;
}

template <typename T, typename M = T>
struct PcgArgs {
  unsigned nblk;
  size_t n;
  // matrix (all rows, sliced) and factors (forward / backward plans, sliced)
  PlanView a_plan;
  GroupView a_gv;
  unsigned a_nsl;
  const M *a_val, *a_dia;
  PlanView f_plan;
  const M *f_val, *f_dia;
  PlanView b_plan;
  const M *b_val;
  // forward level 0 inside the update: per matrix slice, the rows that are
  // level 0 (l0mask) and those that are also backward level 0 (isomask),
  // and their D^-1 blocks packed in the matrix-slice layout
  const unsigned *l0mask, *isomask;
  const M *dinv0;
  const unsigned *fumask;  // per forward slice: rows with no upper blocks
  // vectors
  T *r, *x, *Pr, *p;
  double *partials, *hist;
  double *partials2;  // (r,r) partials of the update kernel
  double *partials3;  // (r,Pr) partials of the fused iteration
  double *partials4;  // mixed: |r_true - r_recursive|^2 partials of a reliable update
  double *partials5;  // mixed: FP64 |b - A x|^2 partials of a reliable update
  unsigned upd_part_off;
  unsigned xcd;  // workgroups of the slice kernels numbered per XCD (wg_id)
  // the SpMV's numbering with xcd set: segments of its grid (seg_wg_id)
  unsigned nseg;
  unsigned seg[kMaxSeg + 1];
  // iteration state, two copies: S is read by the SpMV, the r/x update and
  // the separate beta step, S1 by the sweeps and the folded p update.  The
  // separate scalar-step kernels write both; the p update with the beta step
  // folded in writes S.
  PcgScal *S, *S1;
  // partial counts of the beta step's two sets
  unsigned n_beta2, n_beta3;
  // mixed precision: FP64 solution part (FP32 vectors), right-hand side,
  // matrix and outer history
  double *x64;
  const double *b64;
  const double *a64_val, *a64_dia;
  double *hist64;
};

// x = 0, partial (r,r)
template <typename T, typename M>
__global__ void __launch_bounds__(kVecBlock) k_pcg_init(PcgArgs<T, M> a) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kVecBlock) {
    const double v = (double)a.r[i];
    s += v * v;
    a.x[i] = 0;
  }
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
}

// Scalar steps, one workgroup each, over the partials the previous kernels
// wrote (lsitrsol.h:95-100, 116-117, 122-125, 132-135).
enum ScalMode {
  kScalInit = 0,
  kScalAlpha = 1,
  kScalConvBeta = 2,   // FP64 iteration: |r| history / stop test, then beta
  kScalRPr0 = 3,
  kScalInitMixed = 5,   // |b|^2: outer history[0], rmax2
  kScalAlphaMixed = 6,  // FP64 |b - A x|^2 when a reliable update ran: outer
                        // history, stop test; then alpha
  kScalBetaMixed = 7    // FP32 |r|^2: decide a reliable update for the next
                        // iteration; then beta
};

// One scalar step on a private copy S of the state, from the partial sums t,
// t2, t3; hist entries are written when `writer`.  Every caller with the
// same state and sums computes the same new state.
template <int MODE>
__device__ __attribute__((always_inline)) inline void scal_step(PcgScal &S, double t, double t2,
                                                                double t3, double *hist,
                                                                bool writer) {
  const bool upd = S.upd != 0;  // before this step
  if (MODE == kScalInit) {  // |r0|^2: history[0]; early return when < 1e-30
    if (writer) hist[0] = sqrt(t);
    S.sq = t;
    S.it = 0;
    S.converged = 0;
    S.nan = 0;
    S.done = 0;
    if (t < 1.0e-30) {
      S.done = 1;
      S.converged = 1;
    } else if (!(t == t) || t > 1.0e308) {
      S.done = 1;
      S.nan = 1;
    }
    S.inv_sq0 = 1.0 / t;
  } else if (MODE == kScalAlpha) {  // alpha = rPr / pAp
    S.alpha = S.rPr / t;
  } else if (MODE == kScalRPr0) {
    S.rPr = t;
  } else if (MODE == kScalConvBeta) {
    // push |r|, stop on sqrt(|r|^2/|r0|^2) < tol; else beta = rPr1/rPr
    const int it = S.it + 1;
    S.it = it;
    S.sq = t;
    if (writer) hist[it] = sqrt(t);
    if (sqrt(t * S.inv_sq0) < S.tol) {
      S.done = 1;
      S.converged = 1;
    } else if (!(t == t) || t > 1.0e308) {
      S.done = 1;
      S.nan = 1;
    } else if (it >= S.max_nitr) {
      S.done = 1;
    } else {
      S.beta = t2 / S.rPr;
      S.rPr = t2;
    }
  } else if (MODE == kScalInitMixed) {
    if (writer) hist[0] = sqrt(t);
    S.sq_b = t;
    S.rmax2 = t;
    S.best2 = t;
    S.it_best = 0;
    S.stalled = 0;
    S.sq = t;
    S.it = 0;
    S.nupd = 0;
    S.upd = 0;
    S.final_ = 0;
    S.it_upd = 0;
    S.restart = 0;
    S.converged = 0;
    S.nan = 0;
    S.done = 0;
    if (t < 1.0e-30) {
      S.done = 1;
      S.converged = 1;
    } else if (!(t == t) || t > 1.0e308) {
      S.done = 1;
      S.nan = 1;
    }
  } else if (MODE == kScalAlphaMixed) {
    // t: (p, Ap); t2: FP64 |b - A x|^2 of the reliable update, if one ran
    if (upd) {
      const int k = S.nupd + 1;
      S.nupd = k;
      if (writer) hist[k] = sqrt(t2);
      S.rmax2 = t2;
      if (sqrt(t2 / S.sq_b) < S.tol) {
        S.done = 1;
        S.converged = 1;
      } else if (!(t2 == t2) || t2 > 1.0e308) {
        S.done = 1;
        S.nan = 1;
      } else if (S.final_ || k >= S.max_upd) {
        S.done = 1;
      } else if (t2 < 0.25 * S.best2) {
        S.best2 = t2;
        S.it_best = S.it;
      } else if (S.it - S.it_best >= kStallIters) {
        // the FP32 iteration no longer reduces the FP64 residual (the FP32
        // matrix is too coarse for this system): stop, the host finishes
        // with the FP64 solver
        S.done = 1;
        S.stalled = 1;
      }
    }
    if (upd && (!(t > 0.0) || t3 > 1.0e-2 * t2)) {
      // the recursive residual the current direction was built from is more
      // than 10% off the true one (or reached zero, p = 0): no step now, and
      // the next direction restarts from the true residual
      S.alpha = 0.0;
      S.restart = 1;
    } else {
      S.alpha = S.rPr / t;
    }
  } else {  // kScalBetaMixed: t: FP32 |r|^2 after the update, t2: (r, Pr)
    const int it = S.it + 1;
    S.it = it;
    S.sq = t;
    S.upd = 0;
    if (!(t == t) || t > 1.0e308) {
      S.done = 1;
      S.nan = 1;
      return;
    }
    if (t < S.delta2 * S.rmax2 || sqrt(t / S.sq_b) < S.tol || it >= S.max_nitr ||
        (S.period > 0 && it - S.it_upd >= S.period)) {
      S.upd = 1;
      S.final_ = it >= S.max_nitr ? 1 : 0;
      S.it_upd = it;
    }
    S.beta = S.restart ? 0.0 : t2 / S.rPr;
    S.restart = 0;
    S.rPr = t2;
  }
}

// The fields scal_step<MODE> may change, from s into d (the folded step
// copies the rest unchanged, so the whole state need not stay in registers).
template <int MODE>
__device__ __attribute__((always_inline)) inline void store_step_fields(PcgScal &d,
                                                                        const PcgScal &s) {
  if constexpr (MODE == kScalAlpha) {
    d.alpha = s.alpha;
  } else if constexpr (MODE == kScalAlphaMixed) {
    d.nupd = s.nupd;
    d.rmax2 = s.rmax2;
    d.best2 = s.best2;
    d.it_best = s.it_best;
    d.stalled = s.stalled;
    d.done = s.done;
    d.converged = s.converged;
    d.nan = s.nan;
    d.alpha = s.alpha;
    d.restart = s.restart;
  } else if constexpr (MODE == kScalConvBeta) {
    d.it = s.it;
    d.sq = s.sq;
    d.done = s.done;
    d.converged = s.converged;
    d.nan = s.nan;
    d.beta = s.beta;
    d.rPr = s.rPr;
  } else if constexpr (MODE == kScalBetaMixed) {
    d.it = s.it;
    d.sq = s.sq;
    d.upd = s.upd;
    d.done = s.done;
    d.nan = s.nan;
    d.final_ = s.final_;
    d.it_upd = s.it_upd;
    d.beta = s.beta;
    d.restart = s.restart;
    d.rPr = s.rPr;
  } else {
    copy_state(d, s);
  }
}

// The folded beta step keeps the state in LDS: the first kStateWords threads
// each load one 32-bit word of it early and store it into the block's copy
// (and, in workgroup 0, into the output state: the words the step leaves
// unchanged) before the block's reduction barrier.
constexpr unsigned kStateWords = sizeof(PcgScal) / 4;
static_assert(sizeof(PcgScal) % 4 == 0 && kStateWords <= (unsigned)kWave,
              "the state's words fit one wavefront");

__device__ inline unsigned state_word(const PcgScal *S) {
  return threadIdx.x < kStateWords ? ld_state(reinterpret_cast<const unsigned *>(S) + threadIdx.x)
                                   : 0u;
}

__device__ inline void put_state_word(PcgScal *block_copy, PcgScal *out, unsigned w) {
  if (threadIdx.x < kStateWords) {
    reinterpret_cast<unsigned *>(block_copy)[threadIdx.x] = w;
    if (blockIdx.x == 0) reinterpret_cast<unsigned *>(out)[threadIdx.x] = w;
  }
}

// pa/na: first partial set, pb/nb: second (ConvBeta, AlphaMixed, BetaMixed),
// pc/nc: third (AlphaMixed).  Writes the new state to S[0] and S[1] (the
// copies the iteration reads at its two points; equal here).
template <int MODE>
__global__ void __launch_bounds__(kScalBlock)
k_pcg_scalar(PcgScal *S, const double *__restrict__ pa, unsigned na,
             const double *__restrict__ pb, unsigned nb, const double *__restrict__ pc,
             unsigned nc, double *hist) {
  if (MODE != kScalInit && MODE != kScalInitMixed && ld_state(&S->done)) return;
  const bool upd = ld_state(&S->upd) != 0;  // block-uniform
  const bool two = MODE == kScalConvBeta || MODE == kScalBetaMixed || (MODE == kScalAlphaMixed && upd);
  const bool three = MODE == kScalAlphaMixed && upd;
  const double *const ps[3] = {pa, pb, pc};
  const unsigned ns[3] = {na, two ? nb : 0u, three ? nc : 0u};
  double ts[3];
  reduce_partial_sets<kScalBlock, 3>(ps, ns, ts);
  if (threadIdx.x != 0) return;
  PcgScal st;
  load_state(st, S);
  scal_step<MODE>(st, ts[0], ts[1], ts[2], hist, true);
  copy_state(S[0], st);
  copy_state(S[1], st);
}

// the mixed solver's FP64 solution component j (x64 + x32 for FP32 vectors)
template <typename T, typename M>
__device__ inline double x_full(const PcgArgs<T, M> &a, size_t j) {
  if constexpr (sizeof(T) == 4)
    return a.x64[j] + (double)a.x[j];
  else
    return (double)a.x[j];
}

// Reliable update, row part: r_i = b_i - (A x)_i in FP64 (FP64 matrix
// values, FP64 solution), stored into r; returns r_i^2 and adds the squared
// change of r_i (true minus recursive) to drift.
template <typename T, typename M>
__device__ inline double residual_row(const PcgArgs<T, M> &a, const SliceLane &L, unsigned i,
                                      double &drift) {
  const unsigned K = a.a_plan.len[L.s * kSliceRows + L.r];
  const unsigned q0 = a.a_plan.off[L.s];
  double acc = 0.0;
  for (unsigned k = 0; k < K; ++k) {
    const size_t q = q0 + k;
    const size_t j = kVS * (size_t)a.a_plan.col[q * kSliceRows + L.r];
    const Trip<double> v = load_trip(a.a64_val, q, L.lane);
    acc += v.a * x_full(a, j) + v.b * x_full(a, j + 1) + v.c * x_full(a, j + 2);
  }
  const size_t ii = kVS * (size_t)i;
  acc += dia_row_dot(a.a64_dia, L, x_full(a, ii), x_full(a, ii + 1), x_full(a, ii + 2));
  const double r = a.b64[ii + L.c] - acc;
  const double e = r - (double)a.r[ii + L.c];
  drift += e * e;
  a.r[ii + L.c] = (T)r;
  return r * r;
}

// blocks per load round of the SpMV: 8 with FP32 values and vectors, 4
// when either is FP64 (FP64: 8 or 16 per round measured no faster at 64^3:
// 71.1 / 78.7 us against 71.6 us with 4; FP32 values with FP64 vectors at 8
// took 162 VGPRs, three wavefronts per SIMD)
template <typename T, typename M>
constexpr int spmv_unroll() {
  return sizeof(M) == 4 && sizeof(T) == 4 ? 8 : 4;
}

// Ap -> Pr, partial (p, Ap).  Mixed form with a reliable update due: also
// r = b - A x in FP64 (partials5: |r|^2; partials4: its change), for the
// alpha step that follows.
template <typename T, typename M>
__global__ void __launch_bounds__(kPackBlock) k_pcg_spmv(PcgArgs<T, M> a) {
  if (ld_state(&a.S->done)) return;
  const bool refine = is_mixed<T, M>() && ld_state(&a.S->upd);  // grid-uniform
  const unsigned w = a.xcd ? seg_wg_id(a.seg, a.nseg) : blockIdx.x;
  SliceLane L;
  L.lane = threadIdx.x & (kWave - 1);
  L.s = w * kSlicesPerBlock + threadIdx.x / kWave;
  L.c = L.lane / kSliceRows;
  L.r = L.lane - L.c * kSliceRows;
  const RowSlots R = row_slots(a.a_plan, a.a_gv, L, 0, L.s < a.a_nsl);
  const unsigned i = R.i;
  double part = 0.0, part2 = 0.0, part4 = 0.0;
  if (i != ~0u) {
    // the row's own node and diagonal block first: off the gathers' chain
    const Trip<T> pi = load_node<T, kVS>(a.p, i);
    const Trip<M> d = load_trip<M>(a.a_dia, L.s, L.lane);
    T acc = 0;
    if constexpr (is_mixed<T, M>())  // FP32 copy: difference form, d = row-sum block
      acc = sliced_row_diff_sum<T, M, spmv_unroll<T, M>(), kVS>(
          acc, pi, L, R.len, R.kl, R.kmax, R.off, a.a_plan.col, a.a_val, a.p);
    else
      acc = sliced_row_sum_m<T, M, false, spmv_unroll<T, M>(), kVS>(
          acc, (T)1, L, R.len, R.kl, R.kmax, R.off, a.a_plan.col, a.a_val, a.p);
    acc += (T)d.a * pi.a + (T)d.b * pi.b + (T)d.c * pi.c;  // dia_row_dot_m's products
    a.Pr[kVS * (size_t)i + L.c] = acc;  // Ap lives in Pr (lsitrsol.h:111)
    part = (double)(L.c == 0 ? pi.a : L.c == 1 ? pi.b : pi.c) * (double)acc;
    if constexpr (is_mixed<T, M>())
      if (refine) part2 = residual_row(a, L, i, part4);
  }
  part = block_reduce_sum<kPackBlock>(part);
  if (refine) {
    part2 = block_reduce_sum<kPackBlock>(part2);
    part4 = block_reduce_sum<kPackBlock>(part4);
  }
  if (threadIdx.x == 0) {
    a.partials[w] = part;
    if (refine) {
      a.partials5[w] = part2;
      a.partials4[w] = part4;
    }
  }
}

// Forward level 0 for the level-0 rows of matrix slice L.s, given every
// lane's new r component rv (wavefront-uniform call), the slice's level-0
// and isolated-row masks m / iso and the lane's D^-1 row d (loaded only when
// m is non-zero): Pr_i = D_i^-1 r_i, the same products in the same order as
// the sweep.  Rows that are also backward level 0 add r_i . Pr_i to srp.
template <typename T, typename M>
__device__ inline void fwd_level0(const PcgArgs<T, M> &a, const SliceLane &L, unsigned i, T rv,
                                  unsigned m, unsigned iso, const Trip<M> &d, double &srp) {
  if (!m) return;
  const T t0 = __shfl(rv, (int)L.r, kWave);
  const T t1 = __shfl(rv, (int)L.r + kSliceRows, kWave);
  const T t2 = __shfl(rv, (int)L.r + 2 * kSliceRows, kWave);
  if (i != ~0u && ((m >> L.r) & 1u)) {
    const T z = (T)d.a * t0 + (T)d.b * t1 + (T)d.c * t2;  // dia_row_dot_m's products
    a.Pr[kVS * (size_t)i + L.c] = z;
    if ((iso >> L.r) & 1u) srp += (double)rv * (double)z;
  }
}

// r += -alpha Ap, x += alpha p, partial (r,r); then forward level 0.
// Each wavefront takes kUpdSlices consecutive matrix slices (rows in order,
// components in lanes): all their entries are loaded before any is stored,
// so the slices' loads are in flight together.  Mixed form with FP32
// vectors, right after a reliable update: x32 is first folded into x64.
constexpr unsigned kUpdSlices = 2;

inline unsigned update_grid(unsigned nsl) {
  return std::max(1u, ceil_div(nsl, kSlicesPerBlock * kUpdSlices));
}

template <typename T, typename M>
__global__ void __launch_bounds__(kPackBlock) k_pcg_update(PcgArgs<T, M> a) {
  // the state: before the rows' loads in FP64, after them in the mixed
  // solver (measured at 64^3: mixed solve 59.0 -> 58.7 ms with it after,
  // FP64 103.2 -> 103.4 ms)
  constexpr bool kLate = is_mixed<T, M>();
  double alpha = 0.0;
  bool fold = false;
  if constexpr (!kLate) {
    if (ld_state(&a.S->done)) return;
    alpha = ld_state(&a.S->alpha);
  }
  // no gathers here: the dispatch order serves (an XCD-ordered numbering
  // measured slower on MI355X: FP64 64^3 14.2 -> 16.9 us)
  const unsigned wv = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave;
  const bool iso_rows = a.upd_part_off != ~0u;  // rows with neither lower nor upper blocks exist
  SliceLane L[kUpdSlices];
  bool sl[kUpdSlices];
  unsigned i[kUpdSlices], m[kUpdSlices], iso[kUpdSlices];
  T r0[kUpdSlices], q0[kUpdSlices], x0[kUpdSlices], p0[kUpdSlices];
  Trip<M> dv[kUpdSlices];
  // Both slices' row indices first (a group without the uniform shape loads
  // them: one wait then covers both), then the vector entries of both
  // slices, the slices' level-0 masks as scalar loads (the slice index is
  // wavefront-uniform) and -- once the masks are in, the entries still in
  // flight -- the D^-1 rows of forward level 0: requested with the entries,
  // not after the stores.
#pragma unroll
  for (unsigned h = 0; h < kUpdSlices; ++h) {
    L[h].lane = threadIdx.x & (kWave - 1);
    L[h].s = __builtin_amdgcn_readfirstlane(wv * kUpdSlices + h);
    L[h].c = L[h].lane / kSliceRows;
    L[h].r = L[h].lane - L[h].c * kSliceRows;
    sl[h] = L[h].s < a.a_nsl;
    i[h] = row_of(a.a_plan, a.a_gv, L[h], 0, sl[h]);
  }
#pragma unroll
  for (unsigned h = 0; h < kUpdSlices; ++h) {
    r0[h] = q0[h] = x0[h] = p0[h] = 0;
    if (i[h] != ~0u) {
      const size_t e = kVS * (size_t)i[h] + L[h].c;
      r0[h] = a.r[e];
      q0[h] = a.Pr[e];
      x0[h] = a.x[e];
      p0[h] = a.p[e];
    }
  }
#pragma unroll
  for (unsigned h = 0; h < kUpdSlices; ++h) {
    m[h] = sl[h] ? __ldg(a.l0mask + L[h].s) : 0u;
    iso[h] = sl[h] && iso_rows ? __ldg(a.isomask + L[h].s) : 0u;
  }
#pragma unroll
  for (unsigned h = 0; h < kUpdSlices; ++h) {
    dv[h] = Trip<M>{};
    if (m[h] && L[h].c < 3 && ((m[h] >> L[h].r) & 1u))  // the lanes of level-0 rows
      dv[h] = load_trip<M>(a.dinv0, L[h].s, L[h].lane);
  }
  if constexpr (kLate) {
    if (ld_state(&a.S->done)) return;
    alpha = ld_state(&a.S->alpha);
    fold = sizeof(T) == 4 && ld_state(&a.S->upd);  // an update just ran: fold x32 into x64
  }
  double s2 = 0.0, srp = 0.0;
#pragma unroll
  for (unsigned h = 0; h < kUpdSlices; ++h) {
    T rn = 0;
    if (i[h] != ~0u) {
      const size_t e = kVS * (size_t)i[h] + L[h].c;
      rn = (T)((double)r0[h] + (-alpha) * (double)q0[h]);
      a.r[e] = rn;
      if (fold) {  // x64 += x32, x32 = alpha p
        a.x64[e] += (double)x0[h];
        a.x[e] = (T)(alpha * (double)p0[h]);
      } else {
        a.x[e] = (T)((double)x0[h] + alpha * (double)p0[h]);
      }
      s2 += (double)rn * (double)rn;
    }
    if (sl[h]) fwd_level0(a, L[h], i[h], rn, m[h], iso[h], dv[h], srp);
  }
  s2 = block_reduce_sum<kPackBlock>(s2);
  if (iso_rows) srp = block_reduce_sum<kPackBlock>(srp);
  if (threadIdx.x == 0) {
    const unsigned w = blockIdx.x;
    a.partials2[w] = s2;
    if (iso_rows) a.partials3[a.upd_part_off + w] = srp;
  }
}

// Sweep launch shapes: SPLIT wavefronts per slice (SPLIT = 2: the two
// wavefronts take the even / odd blocks of every row and add their sums
// through LDS, halving each wavefront's chain of dependent loads), U blocks
// per load round.  With SPLIT = 1 the per-component sums run in stored
// block order, exactly as the reference's loops.
template <int SPLIT>
__device__ inline SliceLane sweep_lane(unsigned slice0, bool xcd, unsigned &half,
                                       unsigned &slot) {
  constexpr unsigned SPB = kPackBlock / kWave / SPLIT;  // slices per workgroup
  const unsigned w = threadIdx.x / kWave;
  half = SPLIT == 1 ? 0u : w / SPB;
  slot = w % SPB;
  SliceLane L;
  L.lane = threadIdx.x & (kWave - 1);
  L.s = slice0 + wg_id(xcd) * SPB + slot;
  L.c = L.lane / kSliceRows;
  L.r = L.lane - L.c * kSliceRows;
  return L;
}

// the halves' sums meet in half 0
template <typename T, int SPLIT>
__device__ inline T join_halves(T t, unsigned half, unsigned slot, unsigned lane) {
  if (SPLIT == 1) return t;
  constexpr unsigned SPB = kPackBlock / kWave / SPLIT;
  __shared__ T xs[SPB][kWave];
  if (half == 1) xs[slot][lane] = t;
  __syncthreads();
  return half == 0 ? t + xs[slot][lane] : t;
}

template <int SPLIT>
inline unsigned sweep_grid(unsigned nsl) {
  return std::max(1u, ceil_div(nsl, kPackBlock / kWave / SPLIT));
}

// forward level >= 1: Pr = D^-1 (r - L Pr) on the level's slices (reads r
// directly: the reference's copy Pr = r is fused away).  PART: the level
// holds rows with no upper blocks; their (r, Pr) is summed here.
template <typename T, typename M, bool PART, int SPLIT, int U>
__global__ void __launch_bounds__(kPackBlock)
k_pcg_fwd(PcgArgs<T, M> a, unsigned slice0, unsigned nsl, GroupView gv, double *dst) {
  constexpr bool NT = sweep_nt<M>();
  unsigned half, slot;
  const SliceLane L = sweep_lane<SPLIT>(slice0, a.xcd, half, slot);
  const bool sl = L.s < slice0 + nsl;
  if (SPLIT == 1 && !PART && !sl) return;  // whole wavefront
  const RowSlots R = row_slots(a.f_plan, gv, L, slice0, sl);
  const unsigned i = R.i;
  // r_i, the D_i^-1 row and the first round of blocks are requested before
  // the state is read: off the gathers' chain, and none depends on it
  T rv = 0;
  Trip<M> d{};
  SlotRound<M, U> R0;
  if (i != ~0u) {
    if (half == 0) {
      rv = a.r[kVS * (size_t)i + L.c];
      d = load_trip<M, NT>(a.f_dia, L.s, L.lane);
    }
    load_slot_round<M, U, SPLIT, NT>(R0, L, half, R.kl, R.off, a.f_plan.col, a.f_val);
  }
  if (ld_state(&a.S1->done)) return;
  T t = rv;
  if (i != ~0u)
    t = sliced_row_sum_pre<T, M, true, U, SPLIT, kVS, NT>(t, (T)1, R0, L, half, R.len, R.kl,
                                                          R.kmax, R.off, a.f_plan.col, a.f_val,
                                                          a.Pr);
  t = join_halves<T, SPLIT>(t, half, slot, L.lane);
  double part = 0.0;
  if (half == 0 && sl) {
    const T t0 = __shfl(t, (int)L.r, kWave);
    const T t1 = __shfl(t, (int)L.r + kSliceRows, kWave);
    const T t2 = __shfl(t, (int)L.r + 2 * kSliceRows, kWave);
    if (i != ~0u) {
      const T out = (T)d.a * t0 + (T)d.b * t1 + (T)d.c * t2;  // dia_row_dot_m's products
      a.Pr[kVS * (size_t)i + L.c] = out;
      if (PART && ((a.fumask[L.s] >> L.r) & 1u)) part = (double)rv * (double)out;
    }
  }
  if (PART) {
    part = block_reduce_sum<kPackBlock>(part);
    if (threadIdx.x == 0) dst[wg_id(a.xcd)] = part;
  }
}

// backward level + partial (r, Pr) of the rows it finalises, written at
// dst[workgroup]
template <typename T, typename M, int SPLIT, int U>
__global__ void __launch_bounds__(kPackBlock)
k_pcg_bwd(PcgArgs<T, M> a, unsigned slice0, unsigned nsl, GroupView gv, double *dst) {
  constexpr bool NT = sweep_nt<M>();
  unsigned half, slot;
  const SliceLane L = sweep_lane<SPLIT>(slice0, a.xcd, half, slot);
  const bool sl = L.s < slice0 + nsl;
  const RowSlots R = row_slots(a.b_plan, gv, L, slice0, sl);
  const unsigned i = R.i;
  // the row's own entries and its first round of blocks before the state
  T t = 0, rv = 0;
  SlotRound<M, U> R0;
  if (i != ~0u) {
    if (half == 0) {
      t = a.Pr[kVS * (size_t)i + L.c];
      rv = a.r[kVS * (size_t)i + L.c];
    }
    load_slot_round<M, U, SPLIT, NT>(R0, L, half, R.kl, R.off, a.b_plan.col, a.b_val);
  }
  if (ld_state(&a.S1->done)) return;
  if (i != ~0u)
    t = sliced_row_sum_pre<T, M, true, U, SPLIT, kVS, NT>(t, (T)1, R0, L, half, R.len, R.kl,
                                                          R.kmax, R.off, a.b_plan.col, a.b_val,
                                                          a.Pr);
  t = join_halves<T, SPLIT>(t, half, slot, L.lane);
  double part = 0.0;
  if (half == 0 && i != ~0u) {
    a.Pr[kVS * (size_t)i + L.c] = t;
    part = (double)rv * (double)t;
  }
  part = block_reduce_sum<kPackBlock>(part);
  if (threadIdx.x == 0) dst[wg_id(a.xcd)] = part;
}

// p = beta p + Pr (mode 1), or p = Pr (mode 0, prologue).  Four entries per
// thread with 16-byte loads (FP32; two for FP64), one pass over the grid
// pupdate_grid(n); the remainder of n / 4 by the first threads.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

inline unsigned pupdate_grid(size_t n) { return std::max(1u, ceil_div(n / 4, kVecBlock)); }

template <typename T, typename M>
__global__ void __launch_bounds__(kVecBlock) k_pcg_pupdate(PcgArgs<T, M> a, int mode) {
  if (ld_state(&a.S->done)) return;
  const double beta = ld_state(&a.S->beta);
  using V = typename Vec4<T>::type;
  const size_t k = (size_t)blockIdx.x * kVecBlock + threadIdx.x, n4 = a.n / 4;
  if (k < n4) {
    const V z = reinterpret_cast<const V *>(a.Pr)[k];
    V &pk = reinterpret_cast<V *>(a.p)[k];
    if (mode == 0) {
      pk = z;
    } else {
      const V p = pk;
      pk = V{(T)(beta * (double)p.x + (double)z.x), (T)(beta * (double)p.y + (double)z.y),
             (T)(beta * (double)p.z + (double)z.z), (T)(beta * (double)p.w + (double)z.w)};
    }
  }
  if (k < a.n - 4 * n4) {
    const size_t i = 4 * n4 + k;
    a.p[i] = mode == 0 ? a.Pr[i] : (T)(beta * (double)a.p[i] + (double)a.Pr[i]);
  }
}

// Workgroup size of the p update that runs the beta step itself: the
// scalar kernel's, so that its partial sums come out in the same order.
constexpr int kFoldBlock = kScalBlock;

inline unsigned pupdate_grid_folded(size_t n) { return std::max(1u, ceil_div(n / 4, kFoldBlock)); }

// p = beta p + Pr with the beta step run here: every workgroup sums the
// update's (r,r) partials and the sweeps' (r, Pr) partials itself, with the
// scalar kernel's reduction, and workgroup 0 stores the new state in S (the
// copy the next SpMV reads).  Same products as k_pcg_pupdate's mode 1.
template <typename T, typename M>
__global__ void __launch_bounds__(kFoldBlock) k_pcg_pupdate_folded(PcgArgs<T, M> a) {
  __shared__ PcgScal sst;  // the block's copy of the state
  const unsigned sw = state_word(a.S1);
  using V = typename Vec4<T>::type;
  const size_t k = (size_t)blockIdx.x * kFoldBlock + threadIdx.x, n4 = a.n / 4;
  const bool body = k < n4, tail = k < a.n - 4 * n4;
  V z{}, p{};
  T zt = 0, pt = 0;
  if (body) {
    z = reinterpret_cast<const V *>(a.Pr)[k];
    p = reinterpret_cast<const V *>(a.p)[k];
  }
  if (tail) {
    zt = a.Pr[4 * n4 + k];
    pt = a.p[4 * n4 + k];
  }
  const double *const ps[2] = {a.partials2, a.partials3};
  const unsigned ns[2] = {a.n_beta2, a.n_beta3};
  double s2[2], ts[2];
  partial_sums_thread<kFoldBlock, 2>(ps, ns, s2);
  put_state_word(&sst, a.S, sw);
  partial_sums_block<kFoldBlock, 2>(s2, ts);
  constexpr int kMode = is_mixed<T, M>() ? kScalBetaMixed : kScalConvBeta;
  if (threadIdx.x == 0 && !sst.done) {
    scal_step<kMode>(sst, ts[0], ts[1], 0.0, a.hist, blockIdx.x == 0);
    if (blockIdx.x == 0) store_step_fields<kMode>(*a.S, sst);
  }
  __syncthreads();
  if (sst.done) return;  // block-uniform
  const double beta = sst.beta;
  if (body)
    reinterpret_cast<V *>(a.p)[k] =
        V{(T)(beta * (double)p.x + (double)z.x), (T)(beta * (double)p.y + (double)z.y),
          (T)(beta * (double)p.z + (double)z.z), (T)(beta * (double)p.w + (double)z.w)};
  if (tail) a.p[4 * n4 + k] = (T)(beta * (double)pt + (double)zt);
}

// x = 0 (and x64 = 0 for FP32 vectors), r = b, partial |b|^2 (FP64)
template <typename T, typename M>
__global__ void __launch_bounds__(kVecBlock) k_mixed_init(PcgArgs<T, M> a) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kVecBlock) {
    const double v = a.b64[i];
    s += v * v;
    a.r[i] = (T)v;
    a.x[i] = (T)0;
    if (sizeof(T) == 4) a.x64[i] = 0.0;
  }
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
}

// the mixed solution: x64 + x32 for FP32 vectors, x itself for FP64
template <typename T>
__global__ void k_mixed_out(double *__restrict__ x, const double *__restrict__ x64,
                            const T *__restrict__ xv, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = sizeof(T) == 4 ? x64[i] + (double)xv[i] : (double)xv[i];
}

// ---------------------------------------------------------------------------
// host side

// A captured run of `chunk` fused iterations.
struct ChunkGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> sig;
  unsigned chunk = 0;
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    sig.clear();
  }
};

// iterations per graph launch: large chunks save launches in long solves,
// small ones bound the iterations that still run (as immediate returns)
// after the device has finished, which dominates short solves
constexpr unsigned kChunkLarge = 16, kChunkSmall = 2;

template <typename T>
struct PcgBufs {
  using elem = T;
  DBuf<T> r, x, Pr, p;
  DBuf<double> partials, partials2, partials3, partials4, partials5, hist;
  DBuf<PcgScal> S;
  ChunkGraph g[2];  // iterations per launch: [0] kChunkLarge, [1] kChunkSmall
  void drop_graph() {
    g[0].drop();
    g[1].drop();
  }
  ~PcgBufs() { drop_graph(); }
};

// Per (matrix, factors) pair: the level-0 masks and D^-1 blocks the update
// kernel needs, and where every (r, Pr) partial of an iteration goes.
struct Fusion {
  const void *A = nullptr;
  uint64_t a_serial = ~0ull, p_serial = ~0ull, serial = 0;
  DBuf<unsigned> l0mask, isomask, row_l0;
  std::vector<unsigned> spmv_seg;  // segments of the SpMV grid (seg_wg_id)
  bool any_iso = false;  // rows with neither lower nor upper blocks
  PackedVals<double> d64;
  PackedVals<float> d32;
};

// Where every (r, Pr) partial of a fused iteration goes (partials3), for a
// sweep shape with SPLIT wavefronts per slice.
struct SweepLayout {
  std::vector<unsigned> fwd_off, bwd_off;  // per level (>= 1), ~0u: none
  unsigned upd_off = 0, n3 = 0;
};

static unsigned split_grid(unsigned nsl, int split) {
  return std::max(1u, ceil_div(nsl, kPackBlock / kWave / split));
}

// Launch shapes of the fused iteration: sweep shape {wavefronts per slice,
// blocks per load round} (shapes 0..5) and, for shapes 6..11, the same with
// the slice kernels' workgroups numbered per XCD.  Shape 0 keeps the
// reference's summation order; the XCD numbering changes no sum.
constexpr int kNumSweepBase = 6;
constexpr int kNumSweepShapes = 2 * kNumSweepBase;
constexpr int kSweepSplit[kNumSweepShapes] = {1, 1, 2, 2, 1, 2, 1, 1, 2, 2, 1, 2};
const char *const kSweepNames[kNumSweepShapes] = {
    "split1_u8",     "split1_u4",     "split2_u8",     "split2_u4",
    "split1_u16",    "split2_u16",    "split1_u8_xcd", "split1_u4_xcd",
    "split2_u8_xcd", "split2_u4_xcd", "split1_u16_xcd", "split2_u16_xcd"};
inline unsigned shape_xcd(int shape) { return shape >= kNumSweepBase ? 1u : 0u; }

struct PcgWork {
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_poll[2] = {nullptr, nullptr};
  PcgScal *host_status = nullptr;  // pinned, 2 entries
  PcgBufs<double> f64;
  PcgBufs<float> f32;
  PcgBufs<double> m64;  // mixed solver with FP64 vectors
  DBuf<double> b64, x64, hist64;
  Fusion fu;
  // per solver kind (0 FP64, 1 mixed with FP32 vectors, 2 mixed with FP64
  // vectors): sweep shape (-1: autotune) and the last autotuner timings
  int shape[3] = {-1, -1, -1};
  std::vector<double> tune_ms[3];
  // mixed solver options: Krylov vector precision, forced-update period
  int mixed_vectors64 = 0;
  unsigned mixed_period = 100;
  // iterations per graph launch (the small chunk at the start and near the
  // predicted end, the large one in between)
  unsigned chunk_large = kChunkLarge, chunk_small = kChunkSmall;
  // beta step folded into the p update: -1 by size, 0 never, 1 always
  int fold_mode = -1;
  // the last mixed solve stalled and was finished by the FP64 solver, on
  // these (allocated on the first such solve)
  bool last_mixed_fallback = false;
  DBuf<double> fb_r, fb_e;
  PcgWork() {
    DFM2H_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_poll[0], hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_poll[1], hipEventDisableTiming));
    DFM2H_HIP(hipHostMalloc((void **)&host_status, 2 * sizeof(PcgScal), hipHostMallocDefault));
  }
  ~PcgWork() {
    f64.drop_graph();
    f32.drop_graph();
    m64.drop_graph();
    if (host_status) (void)hipHostFree(host_status);
    for (hipEvent_t e : {ev_in, ev_out, ev_poll[0], ev_poll[1]})
      if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
};

void destroy_pcg_work(PcgWork *w) { delete w; }

static PcgWork &work_of(dfm2h_ilu P) {
  if (!P->work) P->work = new PcgWork;
  return *P->work;
}

static SweepLayout sweep_layout(dfm2h_bsr A, dfm2h_ilu P, int split) {
  SweepLayout Lo;
  unsigned off = 0;
  const size_t nlb = P->bplan.group_ptr.size() - 1, nlf = P->fplan.group_ptr.size() - 1;
  Lo.bwd_off.assign(nlb, ~0u);
  for (size_t l = 1; l < nlb; ++l) {
    Lo.bwd_off[l] = off;
    off += split_grid(P->bplan.group_slices(l), split);
  }
  Lo.fwd_off.assign(nlf, ~0u);
  for (size_t l = 1; l < nlf; ++l)
    if (P->flev_u0[l]) {
      Lo.fwd_off[l] = off;
      off += split_grid(P->fplan.group_slices(l), split);
    }
  if (P->work->fu.any_iso) {  // the update kernel sums those rows' (r, Pr)
    Lo.upd_off = off;
    off += update_grid(A->plan.nslices);
  } else {
    Lo.upd_off = ~0u;
  }
  Lo.n3 = off;
  return Lo;
}

// Builds the fused-iteration data for (A, P) when either pattern changed.
// Matrix slice s holds rows 21 s .. 21 s + 20 (the matrix plan is one group
// of all rows in order).  Row i is forward level 0 iff it has no lower
// blocks, backward level 0 iff it has no upper blocks.
static void ensure_fusion(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W) {
  Fusion &F = W.fu;
  if (F.A == A && F.a_serial == A->pattern_serial && F.p_serial == P->pattern_serial) return;
  const unsigned nsl = A->plan.nslices, n = A->nblk;
  std::vector<unsigned> l0(nsl, 0), iso(nsl, 0), row(std::max<size_t>((size_t)nsl * kSliceRows, 1), ~0u);
  for (unsigned s = 0; s < nsl; ++s)
    for (unsigned r = 0; r < (unsigned)kSliceRows; ++r) {
      const unsigned i = s * kSliceRows + r;
      if (i >= n || P->diaind[i] != P->colind[i]) continue;
      l0[s] |= 1u << r;
      row[i] = i;
      if (P->diaind[i] == P->colind[i + 1]) iso[s] |= 1u << r;
    }
  F.l0mask.upload(l0);
  F.isomask.upload(iso);
  F.any_iso = std::any_of(iso.begin(), iso.end(), [](unsigned m) { return m != 0; });
  F.row_l0.upload(row);
  // SpMV segments: one per forward level when the levels are consecutive row
  // ranges (a colour ordering), boundaries rounded to multiples of 8
  // workgroups; else one segment
  {
    const unsigned gs = slice_grid(nsl);
    const size_t nlev = P->levptr_f.size() - 1;
    const bool ranges = P->flev_ranges && nlev >= 2 && nlev <= kMaxSeg;
    F.spmv_seg.assign(1, 0u);
    if (ranges)
      for (size_t l = 1; l < nlev; ++l) {
        const unsigned wg = (unsigned)(P->levptr_f[l] / (kSliceRows * kSlicesPerBlock));
        const unsigned b = std::min(gs, (wg + kXcds / 2) / kXcds * kXcds);
        if (b > F.spmv_seg.back()) F.spmv_seg.push_back(b);
      }
    if (F.spmv_seg.back() < gs) F.spmv_seg.push_back(gs);
    if (F.spmv_seg.size() < 2) F.spmv_seg.push_back(gs);
  }
  F.d64.version = ~0ull;
  F.d32.version = ~0ull;
  F.A = A;
  F.a_serial = A->pattern_serial;
  F.p_serial = P->pattern_serial;
  F.serial++;
}

// D^-1 of the level-0 rows in matrix-slice layout, precision T
template <typename T>
static void ensure_dinv(dfm2h_bsr A, dfm2h_ilu P, PackedVals<T> &V, hipStream_t s) {
  if (V.version == P->version && (V.dia.p || A->plan.nslices == 0)) return;
  const unsigned nsl = A->plan.nslices;
  V.dia.alloc((size_t)nsl * kSliceVals);
  if (nsl) {
    hipLaunchKernelGGL(k_pack_dia<T>, dim3(pack_grid(nsl)), dim3(256), 0, s, nsl,
                       P->work->fu.row_l0.p, P->vdia.p, V.dia.p);
    check_launch("k_pack_dia");
  }
  V.version = P->version;
}

template <typename T, typename M>
static PcgArgs<T, M> make_args(dfm2h_bsr A, dfm2h_ilu P, PcgBufs<T> &B, const PackedVals<M> &av,
                               const PackedVals<M> &fv, const PackedVals<M> &bv,
                               const PackedVals<M> &dv) {
  const Fusion &F = P->work->fu;
  PcgArgs<T, M> a{};
  a.nblk = A->nblk;
  a.n = (size_t)A->nblk * kVS;
  a.a_plan = view(A->plan);
  a.a_gv = A->plan.gview.empty() ? GroupView{0u, 0u, 0u, 0u, 0u} : A->plan.gview[0];
  a.a_nsl = A->plan.nslices;
  a.a_val = av.val.p;
  a.a_dia = av.dia.p;
  a.f_plan = view(P->fplan);
  a.f_val = fv.val.p;
  a.f_dia = fv.dia.p;
  a.b_plan = view(P->bplan);
  a.b_val = bv.val.p;
  a.l0mask = F.l0mask.p;
  a.isomask = F.isomask.p;
  a.dinv0 = dv.dia.p;
  a.fumask = P->fumask.p;
  a.r = B.r.p;
  a.x = B.x.p;
  a.Pr = B.Pr.p;
  a.p = B.p.p;
  a.partials = B.partials.p;
  a.partials2 = B.partials2.p;
  a.partials3 = B.partials3.p;
  a.partials4 = B.partials4.p;
  a.upd_part_off = 0;  // set per sweep shape
  a.xcd = 0;           // set per launch shape
  a.nseg = (unsigned)F.spmv_seg.size() - 1;
  for (unsigned k = 0; k <= a.nseg; ++k) a.seg[k] = F.spmv_seg[k];
  a.partials5 = B.partials5.p;
  a.hist = B.hist.p;
  a.S = B.S.p;
  a.S1 = B.S.p + 1;
  return a;
}

static unsigned bwd_partial_total(dfm2h_ilu P) {
  unsigned tot = 0;
  for (size_t l = 0; l + 1 < P->bplan.group_ptr.size(); ++l)
    tot += slice_grid(P->bplan.group_slices(l));
  return tot;
}

// Scratch vectors and partial buffers; the vectors are zeroed once here.
template <typename T>
static void ensure_bufs(PcgBufs<T> &B, dfm2h_bsr A, dfm2h_ilu P, unsigned max_nitr) {
  const size_t n = (size_t)A->nblk * kVS;
  if (B.r.n != n) {
    B.r.alloc(n);
    B.x.alloc(n);
    B.Pr.alloc(n);
    B.p.alloc(n);
    B.r.zero();
    B.x.zero();
    B.Pr.zero();
    B.p.zero();
  }
  const size_t gs = slice_grid(A->plan.nslices);
  const size_t np = std::max<size_t>({gs, (size_t)vec_grid(n), (size_t)bwd_partial_total(P), 1});
  if (B.partials.n < np) B.partials.alloc(np);
  if (B.partials2.n < gs) B.partials2.alloc(gs);
  if (B.partials4.n < gs) B.partials4.alloc(gs);
  if (B.partials5.n < gs) B.partials5.alloc(gs);
  const size_t n3 = std::max<size_t>(sweep_layout(A, P, 2).n3, 1);  // the larger layout
  if (B.partials3.n < n3) B.partials3.alloc(n3);
  if (B.hist.n < (size_t)max_nitr + 2) B.hist.alloc((size_t)max_nitr + 2);
  if (!B.S.p) {
    B.S.alloc(2);
    DFM2H_HIP(hipMemsetAsync(B.S.p, 0, 2 * sizeof(PcgScal), stream()));
  }
}

// Prologue form: every forward and backward level, (r,Pr) partials into
// `partials` (one per backward workgroup); returns their count.
template <typename T, typename M>
static unsigned launch_sweeps_full(const PcgArgs<T, M> &a, dfm2h_ilu P, hipStream_t s) {
  for (size_t l = 0; l + 1 < P->fplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->fplan.group_slices(l);
    hipLaunchKernelGGL((k_pcg_fwd<T, M, false, 1, 8>), dim3(slice_grid(nsl)), dim3(kPackBlock), 0,
                       s, a, P->fplan.group_ptr[l], nsl, P->fplan.gview[l], (double *)nullptr);
  }
  unsigned off = 0;
  for (size_t l = 0; l + 1 < P->bplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->bplan.group_slices(l);
    const unsigned g = slice_grid(nsl);
    hipLaunchKernelGGL((k_pcg_bwd<T, M, 1, 8>), dim3(g), dim3(kPackBlock), 0, s, a,
                       P->bplan.group_ptr[l], nsl, P->bplan.gview[l], a.partials + off);
    off += g;
  }
  return off;
}

// Iteration form: forward levels >= 1 then backward levels >= 1 (level 0 of
// each is fused away), (r,Pr) partials into partials3 at layout Lo.
template <typename T, typename M, int SPLIT, int U>
static void launch_sweeps_shape(const PcgArgs<T, M> &a, dfm2h_ilu P, const SweepLayout &Lo,
                                hipStream_t s, int which) {
  for (size_t l = 1; (which & 1) && l + 1 < P->fplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->fplan.group_slices(l);
    const dim3 g(split_grid(nsl, SPLIT));
    if (Lo.fwd_off[l] != ~0u)
      hipLaunchKernelGGL((k_pcg_fwd<T, M, true, SPLIT, U>), g, dim3(kPackBlock), 0, s, a,
                         P->fplan.group_ptr[l], nsl, P->fplan.gview[l],
                         a.partials3 + Lo.fwd_off[l]);
    else
      hipLaunchKernelGGL((k_pcg_fwd<T, M, false, SPLIT, U>), g, dim3(kPackBlock), 0, s, a,
                         P->fplan.group_ptr[l], nsl, P->fplan.gview[l], (double *)nullptr);
  }
  for (size_t l = 1; (which & 2) && l + 1 < P->bplan.group_ptr.size(); ++l) {
    const unsigned nsl = P->bplan.group_slices(l);
    hipLaunchKernelGGL((k_pcg_bwd<T, M, SPLIT, U>), dim3(split_grid(nsl, SPLIT)), dim3(kPackBlock),
                       0, s, a, P->bplan.group_ptr[l], nsl, P->bplan.gview[l],
                       a.partials3 + Lo.bwd_off[l]);
  }
}

// which: 1 forward, 2 backward, 3 both
template <typename T, typename M>
static void launch_sweeps_fused(const PcgArgs<T, M> &a, dfm2h_ilu P, int shape,
                                const SweepLayout &Lo, hipStream_t s, int which = 3) {
  switch (shape % kNumSweepBase) {
    case 0: launch_sweeps_shape<T, M, 1, 8>(a, P, Lo, s, which); break;
    case 1: launch_sweeps_shape<T, M, 1, 4>(a, P, Lo, s, which); break;
    case 2: launch_sweeps_shape<T, M, 2, 8>(a, P, Lo, s, which); break;
    case 3: launch_sweeps_shape<T, M, 2, 4>(a, P, Lo, s, which); break;
    case 4: launch_sweeps_shape<T, M, 1, 16>(a, P, Lo, s, which); break;
    default: launch_sweeps_shape<T, M, 2, 16>(a, P, Lo, s, which); break;
  }
}

// solver kind: 0 FP64, 1 mixed with FP32 vectors, 2 mixed with FP64 vectors
template <typename T, typename M>
constexpr int solver_kind() {
  return sizeof(M) == 8 ? 0 : (sizeof(T) == 4 ? 1 : 2);
}

template <int MODE>
static void launch_scalar(hipStream_t s, PcgScal *S, const double *pa, unsigned na,
                          const double *pb, unsigned nb, double *hist,
                          const double *pc = nullptr, unsigned nc = 0) {
  hipLaunchKernelGGL(k_pcg_scalar<MODE>, dim3(1), dim3(kScalBlock), 0, s, S, pa, na, pb, nb, pc,
                     nc, hist);
}

// The alpha step (a one-workgroup scalar kernel) and the r/x update; the
// beta step (stop test included) and the p update, the former as a separate
// scalar kernel or folded into the latter.  Both forms give the same bits.
template <typename T, typename M>
static void launch_alpha_update(const PcgArgs<T, M> &a, hipStream_t s) {
  const unsigned gs = slice_grid(a.a_nsl);
  if constexpr (is_mixed<T, M>())
    launch_scalar<kScalAlphaMixed>(s, a.S, a.partials, gs, a.partials5, gs, a.hist64,
                                   a.partials4, gs);
  else
    launch_scalar<kScalAlpha>(s, a.S, a.partials, gs, nullptr, 0, a.hist);
  hipLaunchKernelGGL((k_pcg_update<T, M>), dim3(update_grid(a.a_nsl)), dim3(kPackBlock), 0, s, a);
}

template <typename T, typename M>
static void launch_beta_pupdate(const PcgArgs<T, M> &a, bool fold, hipStream_t s) {
  if (fold) {
    hipLaunchKernelGGL((k_pcg_pupdate_folded<T, M>), dim3(pupdate_grid_folded(a.n)),
                       dim3(kFoldBlock), 0, s, a);
    return;
  }
  if constexpr (is_mixed<T, M>())
    launch_scalar<kScalBetaMixed>(s, a.S, a.partials2, a.n_beta2, a.partials3, a.n_beta3,
                                  a.hist);
  else
    launch_scalar<kScalConvBeta>(s, a.S, a.partials2, a.n_beta2, a.partials3, a.n_beta3, a.hist);
  hipLaunchKernelGGL((k_pcg_pupdate<T, M>), dim3(pupdate_grid(a.n)), dim3(kVecBlock), 0, s, a, 1);
}

// One iteration (lsitrsol.h:107-146): spmv, alpha, update (+ forward level
// 0), forward levels >= 1, backward levels >= 1, stop test + beta, p
// update.  The mixed form's beta step also decides whether the next
// iteration starts with a reliable update (the FP32 residual fell by delta
// since the last one, looks converged, or the cap is reached).  That update
// runs inside the next SpMV: r = b - A (x64 + x32) in FP64, tested in FP64
// by the alpha step (stop when |r| / |b| < tol); the r/x update then folds
// x32 into x64 and the same Krylov sequence goes on from the true residual.
template <typename T, typename M>
static void launch_iteration(const PcgArgs<T, M> &a, dfm2h_ilu P, int shape,
                             const SweepLayout &Lo, bool fold, hipStream_t s) {
  hipLaunchKernelGGL((k_pcg_spmv<T, M>), dim3(slice_grid(a.a_nsl)), dim3(kPackBlock), 0, s, a);
  launch_alpha_update(a, s);
  launch_sweeps_fused(a, P, shape, Lo, s);
  launch_beta_pupdate(a, fold, s);
}

// Captures `chunk` iterations into G unless its signature matches.
template <class F>
static void ensure_graph(ChunkGraph &G, hipStream_t s, const std::vector<uintptr_t> &sig,
                         unsigned chunk, F &&one_iteration) {
  if (G.exec && G.sig == sig) return;
  G.drop();
  DFM2H_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (unsigned k = 0; k < chunk; ++k) one_iteration();
  hipError_t e = hipStreamEndCapture(s, &G.graph);
  DFM2H_REQUIRE(e == hipSuccess, DFM2H_ERR_HIP,
                std::string("graph capture: ") + hipGetErrorString(e));
  DFM2H_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
  G.sig = sig;
  G.chunk = chunk;
}

// Captures B's large- and small-chunk graphs of the iteration.
template <typename T, class F>
static void ensure_graphs(const PcgWork &W, PcgBufs<T> &B, hipStream_t s, unsigned max_iters,
                          const std::function<std::vector<uintptr_t>(unsigned)> &sig,
                          F &&one_iteration) {
  const unsigned cl = std::max(1u, std::min(W.chunk_large, max_iters));
  const unsigned cs = std::max(1u, std::min(W.chunk_small, max_iters));
  ensure_graph(B.g[0], s, sig(cl), cl, one_iteration);
  ensure_graph(B.g[1], s, sig(cs), cs, one_iteration);
}

// Launches captured chunks until the device says done (polled one chunk
// behind, so the device never idles) or max_iters iterations were launched;
// returns the status.  Small chunks run first, and again once the rate at
// which |r|^2 has been falling predicts the end within two large chunks of
// the iterations already launched; large chunks in between.  Which graph
// runs changes no result: every kernel reads the device-side state.
template <typename T>
static void run_chunks(PcgWork &W, PcgBufs<T> &B, unsigned max_iters, bool mixed,
                       PcgScal &status) {
  hipStream_t s = W.s;
  bool finished = false, small = true;
  unsigned launched = 0;
  double prev_it = 0.0, prev_l = 0.0;
  bool have_prev = false;
  for (unsigned c = 0; launched < max_iters && !finished; ++c) {
    const ChunkGraph &G = B.g[small ? 1 : 0];
    DFM2H_HIP(hipGraphLaunch(G.exec, s));
    launched += G.chunk;
    DFM2H_HIP(hipMemcpyAsync(&W.host_status[c & 1], B.S.p, sizeof(PcgScal),
                             hipMemcpyDeviceToHost, s));
    DFM2H_HIP(hipEventRecord(W.ev_poll[c & 1], s));
    if (c == 0) continue;
    DFM2H_HIP(hipEventSynchronize(W.ev_poll[(c - 1) & 1]));
    const PcgScal &st = W.host_status[(c - 1) & 1];
    finished = st.done != 0;
    // log |r|^2 of the start (|r0|^2 or |b|^2) and of this state, the target
    const double l_start = mixed ? std::log(st.sq_b) : -std::log(st.inv_sq0);
    const double l_target = l_start + 2.0 * std::log(st.tol);
    const double l = std::log(st.sq);
    double rate = 0.0;  // decrease of log |r|^2 per iteration, recent if possible
    if (have_prev && st.it > prev_it)
      rate = (prev_l - l) / (st.it - prev_it);
    else if (st.it > 0)
      rate = (l_start - l) / st.it;
    if (st.it > prev_it || !have_prev) {
      prev_it = st.it;
      prev_l = l;
      have_prev = true;
    }
    const double in_flight = (double)launched - st.it;
    const double left = rate > 0.0 && std::isfinite(l) ? (l - l_target) / rate - in_flight
                                                       : 1.0e30;
    small = left < 2.0 * B.g[0].chunk;
  }
  DFM2H_HIP(hipMemcpyAsync(&W.host_status[0], B.S.p, sizeof(PcgScal), hipMemcpyDeviceToHost, s));
  DFM2H_HIP(hipStreamSynchronize(s));
  status = W.host_status[0];
}

template <typename T, typename M>
static std::vector<uintptr_t> graph_sig(const PcgArgs<T, M> &a, dfm2h_bsr A, dfm2h_ilu P,
                                        unsigned chunk, bool fold) {
  return {(uintptr_t)a.a_val,    (uintptr_t)a.a_dia,     (uintptr_t)a.f_val,
          (uintptr_t)a.f_dia,    (uintptr_t)a.b_val,     (uintptr_t)a.dinv0,
          (uintptr_t)a.r,        (uintptr_t)a.partials,  (uintptr_t)a.partials2,
          (uintptr_t)a.partials3, (uintptr_t)a.partials4, (uintptr_t)a.hist, (uintptr_t)a.S,
          (uintptr_t)a.a64_val,  (uintptr_t)a.a64_dia,   (uintptr_t)a.x64,
          (uintptr_t)a.b64,      (uintptr_t)a.hist64,    (uintptr_t)chunk,
          (uintptr_t)A,          (uintptr_t)A->pattern_serial, (uintptr_t)P->pattern_serial,
          (uintptr_t)P->work->fu.serial,
          (uintptr_t)P->work->shape[solver_kind<T, M>()], (uintptr_t)fold,
          (uintptr_t)a.partials5};
}

// Autotuner of the launch shape: times kRep rounds of one SpMV plus one
// forward + backward sweep pair for every shape on the current vectors and
// returns the fastest (FP64: of the one-wavefront-per-slice shapes, whose
// sums run in the reference's order).  The SpMV reads r in place of p (p is
// not set yet) and writes Pr and the partials; the sweeps only rewrite Pr;
// the prologue recomputes both afterwards.
template <typename T, typename M>
static int tune_sweeps(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, const PcgArgs<T, M> &a0,
                       std::vector<double> &ms) {
  constexpr int kRep = 20;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DFM2H_HIP(hipEventCreate(&e0));
  DFM2H_HIP(hipEventCreate(&e1));
  ms.assign(kNumSweepShapes, 0.0);
  const unsigned gs = slice_grid(a0.a_nsl);
  // default: split1_u8 with the XCD-ordered numbering, which moves fewer
  // bytes (measured FP64 64^3: SpMV fetch 395 -> 343 MB per launch, sweeps
  // -12%) at equal or lower time
  constexpr int kDefault = kNumSweepBase;
  int best = kDefault;
  for (int i = 0; i < kNumSweepShapes; ++i) {
    const int v = i == 0 ? kDefault : (i <= kDefault ? i - 1 : i);  // the default first
    // FP64 keeps the reference's summation order (one wavefront per slice)
    if (sizeof(M) == 8 && kSweepSplit[v] != 1) continue;
    const SweepLayout Lo = sweep_layout(A, P, kSweepSplit[v]);
    PcgArgs<T, M> a = a0;
    a.upd_part_off = Lo.upd_off;
    a.xcd = shape_xcd(v);
    a.p = a0.r;
    auto round = [&] {
      hipLaunchKernelGGL((k_pcg_spmv<T, M>), dim3(gs), dim3(kPackBlock), 0, W.s, a);
      launch_sweeps_fused(a, P, v, Lo, W.s);
    };
    round();  // warm-up
    DFM2H_HIP(hipEventRecord(e0, W.s));
    for (int k = 0; k < kRep; ++k) round();
    DFM2H_HIP(hipEventRecord(e1, W.s));
    DFM2H_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    DFM2H_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[v] = (double)t / kRep;
    // another shape must beat the default by 3% to replace it
    if (v != kDefault && ms[v] < (best == kDefault ? 0.97 : 1.0) * ms[best]) best = v;
  }
  check_launch("launch-shape autotuner");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// The sweep shape of this solve: the handle's, else the cached one for
// (DOF bucket, precision), else measured now.  Called after the
// prologue's first scalar step; a solve that is already done skips tuning.
template <typename T, typename M>
static int resolve_shape(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, PcgBufs<T> &B,
                         const PcgArgs<T, M> &a) {
  const int p = solver_kind<T, M>();
  if (W.shape[p] >= 0) return W.shape[p];
  const int bucket = dof_bucket(3 * (size_t)A->nblk);
  int v = tune_lookup("pcg", bucket, p, kNumSweepShapes);
  if (v < 0) {
    PcgScal st{};
    DFM2H_HIP(hipMemcpyAsync(&st, B.S.p, sizeof(PcgScal), hipMemcpyDeviceToHost, W.s));
    DFM2H_HIP(hipStreamSynchronize(W.s));
    if (st.done) return 0;
    v = tune_sweeps<T, M>(A, P, W, a, W.tune_ms[p]);
    tune_store("pcg", bucket, p, v, W.tune_ms[p][v]);
  }
  W.shape[p] = v;
  return v;
}

// Redundant partial-sum reads, in bytes per iteration, up to which the beta
// step runs folded into the p update: then every workgroup of the p update
// reads all the partials the step sums, which grows with the square of the
// system size, while the one-workgroup kernel it saves costs the same at
// every size.
constexpr double kFoldBytesMax = 24.0 * (1 << 20);

// The per-shape fields of the iteration's arguments; returns whether the
// beta step runs folded into the p update.
template <typename T, typename M>
static bool iteration_args(PcgArgs<T, M> &a, const PcgWork &W, int shape, const SweepLayout &Lo) {
  a.upd_part_off = Lo.upd_off;
  a.xcd = shape_xcd(shape);
  a.n_beta2 = update_grid(a.a_nsl);
  a.n_beta3 = Lo.n3;
  if (W.fold_mode >= 0) return W.fold_mode != 0;
  const double bytes = 8.0 * (double)pupdate_grid_folded(a.n) * (a.n_beta2 + a.n_beta3);
  return bytes <= kFoldBytesMax;
}

// Runs the fused PCG on B.r (holding b) into B.x; fills status.
template <typename T, typename M>
static void run_pcg(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, PcgBufs<T> &B,
                    const PcgArgs<T, M> &a0, double tol, unsigned max_nitr, PcgScal &status) {
  hipStream_t s = W.s;
  // scalars: tol and cap (a plain copy into the scalar block, not a memset)
  PcgScal init{};
  init.tol = tol;
  init.max_nitr = (int)max_nitr;
  DFM2H_HIP(hipMemcpyAsync(B.S.p, &init, sizeof(PcgScal), hipMemcpyHostToDevice, s));
  const unsigned gv = vec_grid(a0.n);
  hipLaunchKernelGGL((k_pcg_init<T, M>), dim3(gv), dim3(kVecBlock), 0, s, a0);
  launch_scalar<kScalInit>(s, a0.S, a0.partials, gv, nullptr, 0, a0.hist);
  const int shape = resolve_shape(A, P, W, B, a0);
  const SweepLayout Lo = sweep_layout(A, P, kSweepSplit[shape]);
  PcgArgs<T, M> a = a0;
  const bool fold = iteration_args(a, W, shape, Lo);
  const unsigned nb0 = launch_sweeps_full(a, P, s);
  launch_scalar<kScalRPr0>(s, a.S, a.partials, nb0, nullptr, 0, a.hist);
  hipLaunchKernelGGL((k_pcg_pupdate<T, M>), dim3(pupdate_grid(a.n)), dim3(kVecBlock), 0, s, a, 0);
  check_launch("pcg prologue");
  ensure_graphs(W, B, s, max_nitr, [&](unsigned c) { return graph_sig(a, A, P, c, fold); },
                [&] { launch_iteration(a, P, shape, Lo, fold, s); });
  run_chunks(W, B, max_nitr, false, status);
}

template <typename T>
static void run_pcg_mixed(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, PcgBufs<T> &B,
                          const PcgArgs<T, float> &a0, double tol, double delta,
                          unsigned max_outer, unsigned max_inner, PcgScal &status) {
  hipStream_t s = W.s;
  PcgScal init{};
  init.tol = tol;
  init.max_nitr = (int)max_inner;
  init.delta2 = delta * delta;
  init.max_upd = (int)max_outer;
  init.period = (int)W.mixed_period;
  DFM2H_HIP(hipMemcpyAsync(B.S.p, &init, sizeof(PcgScal), hipMemcpyHostToDevice, s));
  const unsigned gv = vec_grid(a0.n);
  hipLaunchKernelGGL((k_mixed_init<T, float>), dim3(gv), dim3(kVecBlock), 0, s, a0);
  launch_scalar<kScalInitMixed>(s, a0.S, a0.partials, gv, nullptr, 0, a0.hist64);
  const int shape = resolve_shape(A, P, W, B, a0);
  const SweepLayout Lo = sweep_layout(A, P, kSweepSplit[shape]);
  PcgArgs<T, float> a = a0;
  const bool fold = iteration_args(a, W, shape, Lo);
  const unsigned nb0 = launch_sweeps_full(a, P, s);
  launch_scalar<kScalRPr0>(s, a.S, a.partials, nb0, nullptr, 0, a.hist);
  hipLaunchKernelGGL((k_pcg_pupdate<T, float>), dim3(pupdate_grid(a.n)), dim3(kVecBlock), 0, s, a,
                     0);
  check_launch("mixed prologue");
  ensure_graphs(W, B, s, max_inner, [&](unsigned c) { return graph_sig(a, A, P, c, fold); },
                [&] { launch_iteration(a, P, shape, Lo, fold, s); });
  // one large chunk beyond the cap: the reliable update that follows it
  run_chunks(W, B, max_inner + B.g[0].chunk, true, status);
}

static void fence_in(PcgWork &W) {
  DFM2H_HIP(hipEventRecord(W.ev_in, stream()));
  DFM2H_HIP(hipStreamWaitEvent(W.s, W.ev_in, 0));
}

// Packed FP32 copies of A's values (in the difference form of
// sliced_row_diff_sum, which k_pcg_spmv uses on them) and of P's factors,
// rounded from FP64
static void ensure_f32(dfm2h_bsr A, dfm2h_ilu P, hipStream_t s) {
  ensure_packed_rowsum(A->plan, A->pk32, A->vcrs.p, A->vdia.p, A->pat.colind_d.p, A->version,
                       s);
  ensure_packed(P->fplan, P->f32, P->vcrs.p, P->vdia.p, P->version, s);
  ensure_packed(P->bplan, P->b32, P->vcrs.p, nullptr, P->version, s);
  ensure_dinv(A, P, P->work->fu.d32, s);
}

static void ensure_f64(dfm2h_bsr A, dfm2h_ilu P, hipStream_t s) {
  ensure_packed(A->plan, A->pk64, A->vcrs.p, A->vdia.p, A->version, s);
  ensure_packed(P->fplan, P->f64, P->vcrs.p, P->vdia.p, P->version, s);
  ensure_packed(P->bplan, P->b64, P->vcrs.p, nullptr, P->version, s);
  ensure_dinv(A, P, P->work->fu.d64, s);
}

}  // namespace dfm2h

using namespace dfm2h;

static void need_solver(dfm2h_bsr A, dfm2h_ilu P) {
  DFM2H_REQUIRE(A && P, DFM2H_ERR_INVALID, "null handle");
  DFM2H_REQUIRE(A->has_pattern && P->has_pattern, DFM2H_ERR_STATE, "pattern not set");
  DFM2H_REQUIRE(A->nblk == P->nblk, DFM2H_ERR_INVALID, "matrix / preconditioner size mismatch");
  flush_merge(A);
}

// mixed solve with Krylov vectors of precision T; returns the status
template <typename T>
static PcgScal mixed_solve(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, PcgBufs<T> &B, const double *b,
                           double *x, double tol, double delta, unsigned max_outer,
                           unsigned max_inner) {
  ensure_bufs(B, A, P, max_inner);
  const size_t nn = A->nblk;
  if (W.b64.n != kVS * nn) {
    W.b64.alloc(kVS * nn);
    W.x64.alloc(kVS * nn);
    W.b64.zero();
    W.x64.zero();
  }
  if (W.hist64.n < (size_t)max_outer + 2) W.hist64.alloc((size_t)max_outer + 2);
  fence_in(W);
  ensure_f32(A, P, W.s);
  ensure_packed(A->plan, A->pk64, A->vcrs.p, A->vdia.p, A->version, W.s);
  hipLaunchKernelGGL((k_pad4<double, double>), dim3(vec_grid(kVS * nn)), dim3(256), 0, W.s,
                     W.b64.p, b, nn);
  PcgArgs<T, float> a = make_args<T, float>(A, P, B, A->pk32, P->f32, P->b32, W.fu.d32);
  a.x64 = W.x64.p;
  a.b64 = W.b64.p;
  a.a64_val = A->pk64.val.p;
  a.a64_dia = A->pk64.dia.p;
  a.hist64 = W.hist64.p;
  PcgScal st{};
  run_pcg_mixed(A, P, W, B, a, tol, delta, max_outer, max_inner, st);
  hipLaunchKernelGGL(k_mixed_out<T>, dim3(vec_grid(3 * nn)), dim3(256), 0, W.s, x, W.x64.p,
                     B.x.p, 3 * nn);
  check_launch("k_mixed_out");
  return st;
}

namespace dfm2h {

// Dot on the calling thread's stream: one partial per workgroup into
// partials[0 .. vec_grid(n)), their fixed-order sum into *res, read back.
// Every partial the sum reads is written first, so nothing needs clearing.
static void dot_into(const double *x, const double *y, size_t n, double *partials, double *res,
                     double *out) {
  const unsigned g = vec_grid(n);
  hipLaunchKernelGGL(k_dot, dim3(g), dim3(kVecBlock), 0, stream(), x, y, n, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kVecBlock), 0, stream(), partials, g, res);
  check_launch("k_dot");
  DFM2H_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, stream()));
  DFM2H_HIP(hipStreamSynchronize(stream()));
}

// Scratch of the allocation-free dot: kVecGridMax workgroup partials, then
// the result.  Each GPU holds its own copy until the process ends; calls on
// one GPU share it.
constexpr size_t kDotScratch = kVecGridMax + 1;
static __device__ __align__(256) double scratch[kDotScratch];

// The device address of `scratch` on the current device, looked up once
// per device (the host-side name is not a device address).
static double *scratch_on_current_device() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double *addr[kMaxDevices] = {};
  int dev = 0;
  DFM2H_HIP(hipGetDevice(&dev));
  DFM2H_REQUIRE(dev >= 0 && dev < kMaxDevices, DFM2H_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    DFM2H_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DFM2H_REQUIRE(p != nullptr, DFM2H_ERR_HIP, "no device address for the dot scratch");
    addr[dev] = static_cast<double *>(p);
  }
  return addr[dev];
}

}  // namespace dfm2h

using namespace dfm2h;

extern "C" {

int dfm2h_vec_dot(const double *x, const double *y, size_t n, double *out) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    if (n == 0) {
      *out = 0.0;
      return;
    }
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(x);
    DFM2H_DEVICE_ARG(y);
    static thread_local DBuf<double> partials, res;
    if (partials.n < kVecGridMax) partials.alloc(kVecGridMax);
    if (!res.p) res.alloc(1);
    dot_into(x, y, n, partials.p, res.p, out);
  });
}

int dfm2h_vec_dot_noalloc(const double *x, const double *y, size_t n, double *out) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    if (n == 0) {
      *out = 0.0;
      return;
    }
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(x);
    DFM2H_DEVICE_ARG(y);
    // the partials of vec_grid(n) workgroups and the result must fit the
    // scratch (vec_grid never exceeds kVecGridMax: any n fits)
    DFM2H_REQUIRE(vec_grid(n) + 1 <= kDotScratch, DFM2H_ERR_INVALID,
                  "too many partial sums for the dot scratch");
    double *sc = scratch_on_current_device();
    dot_into(x, y, n, sc, sc + kVecGridMax, out);
  });
}

int dfm2h_vec_add_scaled(double *y, double alpha, const double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(x);
    DFM2H_DEVICE_ARG(y);
    hipLaunchKernelGGL(k_add_scaled, dim3(vec_grid(n)), dim3(256), 0, stream(), y, alpha, x, n);
    check_launch("k_add_scaled");
  });
}

int dfm2h_vec_scale_and_add(double *y, double beta, const double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(x);
    DFM2H_DEVICE_ARG(y);
    hipLaunchKernelGGL(k_scale_and_add, dim3(vec_grid(n)), dim3(256), 0, stream(), y, beta, x, n);
    check_launch("k_scale_and_add");
  });
}

int dfm2h_vec_copy(double *dst, const double *src, size_t n) {
  return guard(__func__, [&] {
    if (!n || dst == src) return;
    DFM2H_REQUIRE(dst && src, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(dst);
    DFM2H_DEVICE_ARG(src);
    DFM2H_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  });
}

int dfm2h_vec_set_zero(double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(x);
    hipLaunchKernelGGL(k_fill<double>, dim3(vec_grid(n)), dim3(256), 0, stream(), x, 0.0, n);
    check_launch("k_fill");
  });
}

namespace dfm2h {

// The fused FP64 solve behind dfm2h_solve_pcg: r (device) holds b on entry
// and the final residual on return, x receives the solution; h the it + 1
// residual norms.  Synchronises the solver stream.
static PcgScal fp64_solve(dfm2h_bsr A, dfm2h_ilu P, double *r, double *x, double tol,
                          unsigned max_nitr, std::vector<double> &h) {
  PcgWork &W = work_of(P);
  PcgBufs<double> &B = W.f64;
  ensure_fusion(A, P, W);
  ensure_bufs(B, A, P, max_nitr);
  const size_t nn = A->nblk;
  fence_in(W);
  hipLaunchKernelGGL((k_pad4<double, double>), dim3(vec_grid(kVS * nn)), dim3(256), 0, W.s, B.r.p,
                     r, nn);
  ensure_f64(A, P, W.s);
  const PcgArgs<double, double> a =
      make_args<double, double>(A, P, B, A->pk64, P->f64, P->b64, W.fu.d64);
  PcgScal st{};
  run_pcg(A, P, W, B, a, tol, max_nitr, st);
  hipLaunchKernelGGL((k_unpad4<double, double>), dim3(vec_grid(3 * nn)), dim3(256), 0, W.s, r,
                     B.r.p, nn);
  hipLaunchKernelGGL((k_unpad4<double, double>), dim3(vec_grid(3 * nn)), dim3(256), 0, W.s, x,
                     B.x.p, nn);
  check_launch("k_unpad4");
  h.assign((size_t)st.it + 1, 0.0);
  DFM2H_HIP(hipMemcpyAsync(h.data(), B.hist.p, h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, W.s));
  DFM2H_HIP(hipStreamSynchronize(W.s));
  return st;
}

}  // namespace dfm2h

int dfm2h_solve_pcg(dfm2h_bsr A, dfm2h_ilu P, double *r, double *x, double tol,
                    unsigned max_nitr, double *hist, size_t hist_cap, size_t *n_hist) {
  return guard(__func__, [&] {
    need_solver(A, P);
    DFM2H_REQUIRE(r && x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(r);
    DFM2H_DEVICE_ARG(x);
    std::vector<double> h;
    const PcgScal st = fp64_solve(A, P, r, x, tol, max_nitr, h);
    // the reference's final push of the history when not converged
    if (!st.converged && !st.nan) h.push_back(h.back());
    if (hist)
      for (size_t k = 0; k < h.size() && k < hist_cap; ++k) hist[k] = h[k];
    if (n_hist) *n_hist = h.size();
    DFM2H_REQUIRE(!st.nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
  });
}

int dfm2h_solve_pcg_mixed(dfm2h_bsr A, dfm2h_ilu P, const double *b, double *x, double tol,
                          double delta, unsigned max_outer, unsigned max_inner, double *hist,
                          size_t hist_cap, size_t *n_hist, size_t *n_inner) {
  return guard(__func__, [&] {
    need_solver(A, P);
    DFM2H_REQUIRE(b && x, DFM2H_ERR_INVALID, "null vector");
    DFM2H_DEVICE_ARG(b);
    DFM2H_DEVICE_ARG(x);
    DFM2H_REQUIRE(delta > 0.0 && delta < 1.0, DFM2H_ERR_INVALID, "delta must be in (0, 1)");
    DFM2H_REQUIRE(max_outer >= 1 && max_inner >= 1, DFM2H_ERR_INVALID, "caps must be >= 1");
    PcgWork &W = work_of(P);
    ensure_fusion(A, P, W);
    const PcgScal st =
        W.mixed_vectors64
            ? mixed_solve<double>(A, P, W, W.m64, b, x, tol, delta, max_outer, max_inner)
            : mixed_solve<float>(A, P, W, W.f32, b, x, tol, delta, max_outer, max_inner);
    std::vector<double> h((size_t)st.nupd + 1);
    DFM2H_HIP(hipMemcpyAsync(h.data(), W.hist64.p, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, W.s));
    DFM2H_HIP(hipStreamSynchronize(W.s));
    DFM2H_REQUIRE(!st.nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
    size_t inner = (size_t)st.it;
    W.last_mixed_fallback = st.stalled != 0;
    if (st.stalled) {
      // The FP32 iteration stopped reducing the FP64 residual: finish with
      // the FP64 solver on the correction, e = A^-1 (b - A x), x += e.  Its
      // residual norms are those of x + e_k, so they continue the history.
      const size_t n = 3 * (size_t)A->nblk;
      if (W.fb_r.n != n) {
        W.fb_r.alloc(n);
        W.fb_e.alloc(n);
      }
      DFM2H_HIP(hipMemcpyAsync(W.fb_r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
      DFM2H_HIP(hipStreamSynchronize(W.s));
      const auto ok = [](int rc) {
        DFM2H_REQUIRE(rc == DFM2H_OK, rc, dfm2h_last_error());
      };
      ok(dfm2h_bsr_matvec(A, W.fb_r.p, -1.0, x, 1.0));
      double rr = 0.0;
      ok(dfm2h_vec_dot(W.fb_r.p, W.fb_r.p, n, &rr));
      // stop when |b - A (x + e)| / |b| < tol
      const double tol_e = rr > 0.0 ? tol * h[0] / std::sqrt(rr) : 1.0;
      std::vector<double> he;
      const PcgScal se = fp64_solve(A, P, W.fb_r.p, W.fb_e.p, tol_e, max_inner, he);
      DFM2H_REQUIRE(!se.nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
      hipLaunchKernelGGL(k_add_scaled, dim3(vec_grid(n)), dim3(kVecBlock), 0, W.s, x, 1.0,
                         (const double *)W.fb_e.p, n);
      check_launch("k_add_scaled");
      DFM2H_HIP(hipStreamSynchronize(W.s));
      h.insert(h.end(), he.begin(), he.end());
      inner += (size_t)se.it;
    }
    if (hist)
      for (size_t k = 0; k < h.size() && k < hist_cap; ++k) hist[k] = h[k];
    if (n_hist) *n_hist = h.size();
    if (n_inner) *n_inner = inner;
  });
}

int dfm2h_pcg_mixed_fell_back(dfm2h_ilu P, int *fell_back) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P && fell_back, DFM2H_ERR_INVALID, "null argument");
    *fell_back = P->work && P->work->last_mixed_fallback ? 1 : 0;
  });
}

int dfm2h_pcg_set_mixed_options(dfm2h_ilu P, int fp64_vectors, unsigned update_period) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P, DFM2H_ERR_INVALID, "null handle");
    PcgWork &W = work_of(P);
    W.mixed_vectors64 = fp64_vectors ? 1 : 0;
    W.mixed_period = update_period;
  });
}

int dfm2h_pcg_set_scalar_fold(dfm2h_ilu P, int mode) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(mode >= -1 && mode <= 1, DFM2H_ERR_INVALID, "mode must be -1, 0 or 1");
    work_of(P).fold_mode = mode;
  });
}

int dfm2h_pcg_set_graph_chunks(dfm2h_ilu P, unsigned large, unsigned small) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(small >= 1 && small <= large && large <= 1024, DFM2H_ERR_INVALID,
                  "need 1 <= small <= large <= 1024");
    PcgWork &W = work_of(P);
    W.chunk_large = large;
    W.chunk_small = small;
  });
}

// the solver kind a `fp32` argument names: 0 FP64, else the mixed solver
// with its current vector precision
static int kind_of(PcgWork &W, int fp32) { return fp32 ? (W.mixed_vectors64 ? 2 : 1) : 0; }

int dfm2h_pcg_set_shape(dfm2h_ilu P, int fp32, int shape) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P, DFM2H_ERR_INVALID, "null handle");
    DFM2H_REQUIRE(shape >= -1 && shape < kNumSweepShapes, DFM2H_ERR_INVALID, "bad shape");
    PcgWork &W = work_of(P);
    W.shape[kind_of(W, fp32)] = shape;
    W.f64.drop_graph();
    W.f32.drop_graph();
    W.m64.drop_graph();
  });
}

int dfm2h_pcg_get_shape(dfm2h_ilu P, int fp32, int *shape, const char **name, double *tune_ms,
                        size_t cap, size_t *n_shapes) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(P && shape, DFM2H_ERR_INVALID, "null argument");
    PcgWork &W = work_of(P);
    const int p = kind_of(W, fp32);
    *shape = W.shape[p];
    if (name) *name = W.shape[p] >= 0 ? kSweepNames[W.shape[p]] : "autotune";
    for (size_t k = 0; tune_ms && k < cap && k < W.tune_ms[p].size(); ++k) tune_ms[k] = W.tune_ms[p][k];
    if (n_shapes) *n_shapes = kNumSweepShapes;
  });
}

int dfm2h_pcg_kernel_times(dfm2h_bsr A, dfm2h_ilu P, int fp32, unsigned nrep, double *ms) {
  return guard(__func__, [&] {
    need_solver(A, P);
    DFM2H_REQUIRE(ms && nrep > 0, DFM2H_ERR_INVALID, "bad argument");
    PcgWork &W = work_of(P);
    // Runs nrep iterations with direct launches and an event between
    // consecutive kernel classes, so every class is timed inside the real
    // dependency chain (as in the captured graph).  Convergence and
    // refinement are disabled for the duration.
    auto run = [&](auto tag, auto mtag, auto &B, const auto &a0) {
      using T = decltype(tag);
      using M = decltype(mtag);
      const int shape = std::max(0, W.shape[solver_kind<T, M>()]);
      const SweepLayout Lo = sweep_layout(A, P, kSweepSplit[shape]);
      auto a = a0;
      const bool fold = iteration_args(a, W, shape, Lo);
      PcgScal st{};
      DFM2H_HIP(hipMemcpyAsync(&st, B.S.p, sizeof(PcgScal), hipMemcpyDeviceToHost, W.s));
      DFM2H_HIP(hipStreamSynchronize(W.s));
      PcgScal tst[2] = {st, st};
      for (PcgScal &t : tst) {
        t.done = 0;
        t.upd = 0;
        t.tol = 0.0;
        t.delta2 = 0.0;
        t.max_nitr = 1 << 30;
      }
      DFM2H_HIP(hipMemcpyAsync(B.S.p, tst, 2 * sizeof(PcgScal), hipMemcpyHostToDevice, W.s));
      const unsigned gs = slice_grid(a.a_nsl);
      // segments: spmv | alpha step | update | (two empty) | fwd | bwd |
      // beta step + p update
      constexpr int kSeg = 8;
      TimingEvents<kSeg + 1> ev;
      double acc[kSeg] = {0};
      for (unsigned rep = 0; rep <= nrep; ++rep) {
        DFM2H_HIP(hipEventRecord(ev[0], W.s));
        hipLaunchKernelGGL((k_pcg_spmv<T, M>), dim3(gs), dim3(kPackBlock), 0, W.s, a);
        DFM2H_HIP(hipEventRecord(ev[1], W.s));
        if constexpr (is_mixed<T, M>())
          launch_scalar<kScalAlphaMixed>(W.s, a.S, a.partials, gs, a.partials5, gs, a.hist64,
                                         a.partials4, gs);
        else
          launch_scalar<kScalAlpha>(W.s, a.S, a.partials, gs, nullptr, 0, a.hist);
        DFM2H_HIP(hipEventRecord(ev[2], W.s));
        hipLaunchKernelGGL((k_pcg_update<T, M>), dim3(update_grid(a.a_nsl)), dim3(kPackBlock), 0,
                           W.s, a);
        DFM2H_HIP(hipEventRecord(ev[3], W.s));
        DFM2H_HIP(hipEventRecord(ev[4], W.s));
        DFM2H_HIP(hipEventRecord(ev[5], W.s));
        launch_sweeps_fused(a, P, shape, Lo, W.s, 1);
        DFM2H_HIP(hipEventRecord(ev[6], W.s));
        launch_sweeps_fused(a, P, shape, Lo, W.s, 2);
        DFM2H_HIP(hipEventRecord(ev[7], W.s));
        launch_beta_pupdate(a, fold, W.s);
        DFM2H_HIP(hipEventRecord(ev[8], W.s));
        DFM2H_HIP(hipEventSynchronize(ev[8]));
        if (rep == 0) continue;  // warm-up iteration
        for (int k = 0; k < kSeg; ++k) {
          float t = 0.f;
          DFM2H_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
          acc[k] += t;
        }
      }
      // the SpMV alone, nrep launches back to back between two events: its
      // duration plus the dispatch gap, without the events between classes
      DFM2H_HIP(hipEventRecord(ev[0], W.s));
      for (unsigned rep = 0; rep < nrep; ++rep)
        hipLaunchKernelGGL((k_pcg_spmv<T, M>), dim3(gs), dim3(kPackBlock), 0, W.s, a);
      DFM2H_HIP(hipEventRecord(ev[1], W.s));
      DFM2H_HIP(hipEventSynchronize(ev[1]));
      float t_b2b = 0.f;
      DFM2H_HIP(hipEventElapsedTime(&t_b2b, ev[0], ev[1]));
      check_launch("kernel timing");
      ms[0] = acc[0] / nrep;  // SpMV
      ms[1] = acc[2] / nrep;  // update + forward level 0
      ms[2] = acc[5] / nrep;  // forward levels >= 1
      ms[3] = acc[6] / nrep;  // backward levels >= 1
      ms[4] = acc[7] / nrep;  // (stop test +) beta step + p update
      ms[5] = acc[1] / nrep;  // alpha step
      ms[6] = (double)t_b2b / nrep;  // SpMV, back to back
      ms[7] = fold ? 1.0 : 0.0;  // the beta step ran folded into the p update
      st.done = 1;
      const PcgScal fin[2] = {st, st};
      DFM2H_HIP(hipMemcpyAsync(B.S.p, fin, 2 * sizeof(PcgScal), hipMemcpyHostToDevice, W.s));
      DFM2H_HIP(hipStreamSynchronize(W.s));
    };
    if (fp32) {
      DFM2H_REQUIRE(A->pk32.val.p && W.x64.p && W.fu.A == A, DFM2H_ERR_STATE,
                    "run a mixed solve first");
      auto mixed_args = [&](auto &B) {
        using T = typename std::remove_reference_t<decltype(B)>::elem;
        PcgArgs<T, float> a = make_args<T, float>(A, P, B, A->pk32, P->f32, P->b32, W.fu.d32);
        a.x64 = W.x64.p;
        a.b64 = W.b64.p;
        a.a64_val = A->pk64.val.p;
        a.a64_dia = A->pk64.dia.p;
        a.hist64 = W.hist64.p;
        return a;
      };
      if (W.mixed_vectors64) {
        DFM2H_REQUIRE(W.m64.S.p, DFM2H_ERR_STATE, "run a mixed solve first");
        run(double{}, float{}, W.m64, mixed_args(W.m64));
      } else {
        DFM2H_REQUIRE(W.f32.S.p, DFM2H_ERR_STATE, "run a mixed solve first");
        run(float{}, float{}, W.f32, mixed_args(W.f32));
      }
    } else {
      DFM2H_REQUIRE(W.f64.S.p && A->pk64.val.p && W.fu.A == A, DFM2H_ERR_STATE,
                    "run an FP64 solve first");
      PcgArgs<double, double> a =
          make_args<double, double>(A, P, W.f64, A->pk64, P->f64, P->b64, W.fu.d64);
      run(double{}, double{}, W.f64, a);
    }
  });
}

}  // extern "C"
