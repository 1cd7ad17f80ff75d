// pcg.hip -- view_vectorx.h operator set and the fused Solve_PCG
// (lsitrsol.h:80-148) in FP64, plus FP32-inner / FP64-refinement.
//
// Fused iteration (one HIP graph node per kernel, K iterations per graph):
//   spmv   Ap = A p, partial (p,Ap); last block: alpha = rPr/pAp
//   update r += -alpha Ap, x += alpha p, partial (r,r); last block: history,
//          convergence test, iteration count
//   fwd    one kernel per dependency level: Pr = L^-1 r (reads r directly,
//          which fuses the reference's copy Pr = r)
//   bwd    one kernel per level: Pr = U^-1 Pr, partial (r,Pr); the last
//          block of the last level: beta = rPr1/rPr, rPr = rPr1
//   pupd   p = beta p + Pr
// Every kernel returns at once when the device-side `done` flag is set, so
// iterations launched after convergence cost only their launch.  Scalars
// never leave HBM; the host polls `done` once per chunk, one chunk behind.
// All grid reductions are single-pass with a fixed partition and a fixed
// final order, hence deterministic run to run.
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace dfm2h {

struct PcgScal {
  double rPr;
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
  unsigned ticket[4];
};

constexpr int kVecBlock = 256;
constexpr unsigned kVecGridMax = 2048;

static unsigned vec_grid(size_t n) {
  return std::max(1u, std::min(kVecGridMax, ceil_div(n, kVecBlock)));
}

// ---------------------------------------------------------------------------
// generic deterministic dot (host-facing Dot)

__global__ void __launch_bounds__(kVecBlock)
k_dot(const double *__restrict__ x, const double *__restrict__ y, size_t n,
      double *partials, unsigned *ticket, double *out) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kVecBlock)
    s += x[i] * y[i];
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
  if (last_block(ticket, gridDim.x)) {
    const double t = reduce_partials<kVecBlock>(partials, gridDim.x);
    if (threadIdx.x == 0) {
      *out = t;
      *ticket = 0;
    }
  }
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

// ---------------------------------------------------------------------------
// fused PCG kernels

template <typename T>
struct PcgArgs {
  unsigned nblk;
  size_t n;
  // matrix
  const unsigned *a_colind, *a_rowptr;
  const T *a_vcrs, *a_vdia;
  // factors
  const unsigned *l_colind, *l_rowptr, *l_diaind;
  const T *l_vcrs, *l_vdia;
  // vectors
  T *r, *x, *Pr, *p;
  double *partials, *hist;
  PcgScal *S;
};

// x = 0, sq0 = (r,r), history[0]; stop at once when sq0 < 1e-30
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_pcg_init(PcgArgs<T> a) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kVecBlock) {
    const double v = (double)a.r[i];
    s += v * v;
    a.x[i] = 0;
  }
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  if (last_block(&a.S->ticket[0], gridDim.x)) {
    const double sq0 = reduce_partials<kVecBlock>(a.partials, gridDim.x);
    if (threadIdx.x == 0) {
      a.S->ticket[0] = 0;
      a.hist[0] = sqrt(sq0);
      a.S->sq = sq0;
      a.S->it = 0;
      a.S->converged = 0;
      a.S->nan = 0;
      if (sq0 < 1.0e-30) {
        a.S->done = 1;
        a.S->converged = 1;
      } else if (!(sq0 == sq0) || sq0 > 1.0e308) {
        a.S->done = 1;
        a.S->nan = 1;
      } else {
        a.S->done = 0;
      }
      a.S->inv_sq0 = 1.0 / sq0;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(PcgArgs<T> a) {
  if (a.S->done) return;
  const RowLane L = row_lane();
  const bool valid = L.r < kRowsPerWave && L.slot < a.nblk;
  double part = 0.0;
  if (valid) {
    const unsigned i = L.slot, c = L.c;
    T acc = 0;
    const unsigned s1 = a.a_colind[i + 1];
    for (unsigned s = a.a_colind[i]; s < s1; ++s) {
      const size_t j = a.a_rowptr[s];
      acc += blk_row_dot(a.a_vcrs + (size_t)s * 9 + 3 * c, a.p + 3 * j);
    }
    acc += blk_row_dot(a.a_vdia + (size_t)i * 9 + 3 * c, a.p + 3 * (size_t)i);
    a.Pr[3 * (size_t)i + c] = acc;  // Ap lives in Pr (lsitrsol.h:111)
    part = (double)a.p[3 * (size_t)i + c] * (double)acc;
  }
  part = block_reduce_sum<kBlock>(part);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = part;
  if (last_block(&a.S->ticket[1], gridDim.x)) {
    const double pAp = reduce_partials<kBlock>(a.partials, gridDim.x);
    if (threadIdx.x == 0) {
      a.S->ticket[1] = 0;
      a.S->alpha = a.S->rPr / pAp;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_pcg_update(PcgArgs<T> a) {
  if (a.S->done) return;
  const double alpha = a.S->alpha;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kVecBlock) {
    const T rn = (T)((double)a.r[i] + (-alpha) * (double)a.Pr[i]);
    a.r[i] = rn;
    a.x[i] = (T)((double)a.x[i] + alpha * (double)a.p[i]);
    s += (double)rn * (double)rn;
  }
  s = block_reduce_sum<kVecBlock>(s);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  if (last_block(&a.S->ticket[2], gridDim.x)) {
    const double sq = reduce_partials<kVecBlock>(a.partials, gridDim.x);
    if (threadIdx.x == 0) {
      a.S->ticket[2] = 0;
      const int it = a.S->it + 1;
      a.S->it = it;
      a.S->sq = sq;
      a.hist[it] = sqrt(sq);
      if (sqrt(sq * a.S->inv_sq0) < a.S->tol) {
        a.S->done = 1;
        a.S->converged = 1;
      } else if (!(sq == sq) || sq > 1.0e308) {
        a.S->done = 1;
        a.S->nan = 1;
      } else if (it >= a.S->max_nitr) {
        a.S->done = 1;
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_pcg_fwd_level(PcgArgs<T> a, const unsigned *__restrict__ rows, unsigned nrows) {
  if (a.S->done) return;
  const RowLane L = row_lane();
  const bool valid = L.r < kRowsPerWave && L.slot < nrows;
  const unsigned i = valid ? rows[L.slot] : 0u;
  fwd_row(i, L.c, valid, a.l_colind, a.l_rowptr, a.l_diaind, a.l_vcrs, a.l_vdia,
          (const T *)a.r, a.Pr);
}

// backward level + partial (r, Pr) of the rows it finalises.  The last
// level reduces all levels' partials: prologue (mode 0) sets rPr, an
// iteration (mode 1) sets beta = rPr1/rPr and rPr = rPr1.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_pcg_bwd_level(PcgArgs<T> a, const unsigned *__restrict__ rows, unsigned nrows,
                unsigned part_off, unsigned part_total, int last_level, int mode) {
  if (a.S->done) return;
  const RowLane L = row_lane();
  const bool valid = L.r < kRowsPerWave && L.slot < nrows;
  double part = 0.0;
  if (valid) {
    const unsigned i = rows[L.slot];
    const T t = bwd_row(i, L.c, a.l_colind, a.l_rowptr, a.l_diaind, a.l_vcrs, a.Pr);
    part = (double)a.r[3 * (size_t)i + L.c] * (double)t;
  }
  part = block_reduce_sum<kBlock>(part);
  if (threadIdx.x == 0) a.partials[part_off + blockIdx.x] = part;
  if (!last_level) return;
  if (last_block(&a.S->ticket[3], gridDim.x)) {
    const double rPr1 = reduce_partials<kBlock>(a.partials, part_total);
    if (threadIdx.x == 0) {
      a.S->ticket[3] = 0;
      if (mode == 0) {
        a.S->rPr = rPr1;
      } else {
        a.S->beta = rPr1 / a.S->rPr;
        a.S->rPr = rPr1;
      }
    }
  }
}

// p = beta p + Pr (mode 1), or p = Pr (mode 0, prologue)
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_pcg_pupdate(PcgArgs<T> a, int mode) {
  if (a.S->done) return;
  const double beta = a.S->beta;
  for (size_t i = (size_t)blockIdx.x * kVecBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kVecBlock) {
    if (mode == 0)
      a.p[i] = a.Pr[i];
    else
      a.p[i] = (T)(beta * (double)a.p[i] + (double)a.Pr[i]);
  }
}

// r = b - A x (FP64) written as FP32 for the inner solve, and |r|^2
__global__ void __launch_bounds__(kBlock)
k_residual_to_f32(unsigned nblk, const unsigned *__restrict__ colind,
                  const unsigned *__restrict__ rowptr, const double *__restrict__ vcrs,
                  const double *__restrict__ vdia, const double *__restrict__ b,
                  const double *__restrict__ x, float *__restrict__ r32, double *partials,
                  unsigned *ticket, double *sq_out) {
  const RowLane L = row_lane();
  const bool valid = L.r < kRowsPerWave && L.slot < nblk;
  double part = 0.0;
  if (valid) {
    const unsigned i = L.slot, c = L.c;
    double acc = 0.0;
    const unsigned s1 = colind[i + 1];
    for (unsigned s = colind[i]; s < s1; ++s) {
      const size_t j = rowptr[s];
      acc += blk_row_dot(vcrs + (size_t)s * 9 + 3 * c, x + 3 * j);
    }
    acc += blk_row_dot(vdia + (size_t)i * 9 + 3 * c, x + 3 * (size_t)i);
    const double r = b[3 * (size_t)i + c] - acc;
    r32[3 * (size_t)i + c] = (float)r;
    part = r * r;
  }
  part = block_reduce_sum<kBlock>(part);
  if (threadIdx.x == 0) partials[blockIdx.x] = part;
  if (last_block(ticket, gridDim.x)) {
    const double t = reduce_partials<kBlock>(partials, gridDim.x);
    if (threadIdx.x == 0) {
      *sq_out = t;
      *ticket = 0;
    }
  }
}

__global__ void k_add_f32_to_f64(double *__restrict__ x, const float *__restrict__ d, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] += (double)d[i];
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
struct PcgBufs {
  DBuf<T> r, x, Pr, p;
  DBuf<double> partials, hist;
  DBuf<PcgScal> S;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> sig;
  unsigned chunk = 0;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    sig.clear();
  }
  ~PcgBufs() { drop_graph(); }
};

struct PcgWork {
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_poll[2] = {nullptr, nullptr};
  PcgScal *host_status = nullptr;  // pinned, 2 entries
  PcgBufs<double> f64;
  PcgBufs<float> f32;
  DBuf<double> b64, sq, mixed_partials;
  DBuf<unsigned> ticket;
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

template <typename T>
static PcgArgs<T> make_args(dfm2h_bsr A, dfm2h_ilu P, PcgBufs<T> &B, const T *a_vcrs,
                            const T *a_vdia, const T *l_vcrs, const T *l_vdia) {
  PcgArgs<T> a;
  a.nblk = A->nblk;
  a.n = (size_t)A->nblk * 3;
  a.a_colind = A->pat.colind_d.p;
  a.a_rowptr = A->pat.rowptr_d.p;
  a.a_vcrs = a_vcrs;
  a.a_vdia = a_vdia;
  a.l_colind = P->colind_d.p;
  a.l_rowptr = P->rowptr_d.p;
  a.l_diaind = P->diaind_d.p;
  a.l_vcrs = l_vcrs;
  a.l_vdia = l_vdia;
  a.r = B.r.p;
  a.x = B.x.p;
  a.Pr = B.Pr.p;
  a.p = B.p.p;
  a.partials = B.partials.p;
  a.hist = B.hist.p;
  a.S = B.S.p;
  return a;
}

template <typename T>
static unsigned bwd_partial_total(dfm2h_ilu P) {
  unsigned tot = 0;
  for (size_t l = 0; l + 1 < P->levptr_b.size(); ++l)
    tot += std::max(1u, ceil_div(P->levptr_b[l + 1] - P->levptr_b[l], kRowsPerBlock));
  return tot;
}

template <typename T>
static void ensure_bufs(PcgBufs<T> &B, dfm2h_bsr A, dfm2h_ilu P, unsigned max_nitr) {
  const size_t n = (size_t)A->nblk * 3;
  B.r.alloc(n);
  B.x.alloc(n);
  B.Pr.alloc(n);
  B.p.alloc(n);
  const size_t np = std::max<size_t>({(size_t)ceil_div(A->nblk, kRowsPerBlock),
                                      (size_t)vec_grid(n), (size_t)bwd_partial_total<T>(P), 1});
  if (B.partials.n < np) B.partials.alloc(np);
  if (B.hist.n < (size_t)max_nitr + 2) B.hist.alloc((size_t)max_nitr + 2);
  if (!B.S.p) {
    B.S.alloc(1);
    DFM2H_HIP(hipMemsetAsync(B.S.p, 0, sizeof(PcgScal), stream()));
  }
}

template <typename T>
static void launch_sweeps(const PcgArgs<T> &a, dfm2h_ilu P, hipStream_t s, int mode) {
  for (size_t l = 0; l + 1 < P->levptr_f.size(); ++l) {
    const unsigned nrows = P->levptr_f[l + 1] - P->levptr_f[l];
    hipLaunchKernelGGL(k_pcg_fwd_level<T>, dim3(std::max(1u, ceil_div(nrows, kRowsPerBlock))),
                       dim3(kBlock), 0, s, a, P->levrows_f.p + P->levptr_f[l], nrows);
  }
  const unsigned tot = bwd_partial_total<T>(P);
  unsigned off = 0;
  const size_t nl = P->levptr_b.size() - 1;
  for (size_t l = 0; l < nl; ++l) {
    const unsigned nrows = P->levptr_b[l + 1] - P->levptr_b[l];
    const unsigned g = std::max(1u, ceil_div(nrows, kRowsPerBlock));
    hipLaunchKernelGGL(k_pcg_bwd_level<T>, dim3(g), dim3(kBlock), 0, s, a,
                       P->levrows_b.p + P->levptr_b[l], nrows, off, tot, (int)(l + 1 == nl), mode);
    off += g;
  }
}

template <typename T>
static void launch_iteration(const PcgArgs<T> &a, dfm2h_ilu P, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_spmv<T>, dim3(std::max(1u, ceil_div(a.nblk, kRowsPerBlock))),
                     dim3(kBlock), 0, s, a);
  hipLaunchKernelGGL(k_pcg_update<T>, dim3(vec_grid(a.n)), dim3(kVecBlock), 0, s, a);
  launch_sweeps(a, P, s, 1);
  hipLaunchKernelGGL(k_pcg_pupdate<T>, dim3(vec_grid(a.n)), dim3(kVecBlock), 0, s, a, 1);
}

// Runs the fused PCG on B.r (holding b) into B.x.  Returns iterations done;
// fills status.
template <typename T>
static void run_pcg(dfm2h_bsr A, dfm2h_ilu P, PcgWork &W, PcgBufs<T> &B, const PcgArgs<T> &a,
                    double tol, unsigned max_nitr, PcgScal &status) {
  hipStream_t s = W.s;
  // scalars: tol and cap (a plain copy into the scalar block, not a memset)
  PcgScal init{};
  init.tol = tol;
  init.max_nitr = (int)max_nitr;
  DFM2H_HIP(hipMemcpyAsync(B.S.p, &init, sizeof(PcgScal), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pcg_init<T>, dim3(vec_grid(a.n)), dim3(kVecBlock), 0, s, a);
  launch_sweeps(a, P, s, 0);
  hipLaunchKernelGGL(k_pcg_pupdate<T>, dim3(vec_grid(a.n)), dim3(kVecBlock), 0, s, a, 0);
  check_launch("pcg prologue");

  // graph of `chunk` iterations, recaptured only when buffers or schedule change
  const unsigned chunk = std::max(1u, std::min(16u, max_nitr));
  std::vector<uintptr_t> sig = {(uintptr_t)a.a_vcrs, (uintptr_t)a.a_vdia, (uintptr_t)a.l_vcrs,
                                (uintptr_t)a.l_vdia, (uintptr_t)a.r,      (uintptr_t)a.partials,
                                (uintptr_t)a.hist,   (uintptr_t)a.S,      (uintptr_t)P->levrows_f.p,
                                (uintptr_t)P->levrows_b.p, (uintptr_t)chunk,
                                (uintptr_t)P->levptr_f.size(), (uintptr_t)P->levptr_b.size(),
                                (uintptr_t)a.nblk};
  if (!B.exec || B.sig != sig) {
    B.drop_graph();
    DFM2H_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (unsigned k = 0; k < chunk; ++k) launch_iteration(a, P, s);
    hipError_t e = hipStreamEndCapture(s, &B.graph);
    DFM2H_REQUIRE(e == hipSuccess, DFM2H_ERR_HIP,
                  std::string("graph capture: ") + hipGetErrorString(e));
    DFM2H_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    B.sig = sig;
    B.chunk = chunk;
  }
  const unsigned nchunks = (max_nitr + chunk - 1) / chunk;
  unsigned launched = 0;
  bool finished = false;
  for (unsigned c = 0; c < nchunks && !finished; ++c) {
    DFM2H_HIP(hipGraphLaunch(B.exec, s));
    DFM2H_HIP(hipMemcpyAsync(&W.host_status[c & 1], B.S.p, sizeof(PcgScal),
                             hipMemcpyDeviceToHost, s));
    DFM2H_HIP(hipEventRecord(W.ev_poll[c & 1], s));
    ++launched;
    if (c > 0) {  // poll one chunk behind so the device never idles
      DFM2H_HIP(hipEventSynchronize(W.ev_poll[(c - 1) & 1]));
      finished = W.host_status[(c - 1) & 1].done != 0;
    }
  }
  DFM2H_HIP(hipMemcpyAsync(&W.host_status[0], B.S.p, sizeof(PcgScal), hipMemcpyDeviceToHost, s));
  DFM2H_HIP(hipStreamSynchronize(s));
  status = W.host_status[0];
  (void)launched;
}

static void fence_in(PcgWork &W) {
  DFM2H_HIP(hipEventRecord(W.ev_in, stream()));
  DFM2H_HIP(hipStreamWaitEvent(W.s, W.ev_in, 0));
}

// FP32 copies of A's values and P's factors, refreshed when either changed
static void ensure_f32(dfm2h_bsr A, dfm2h_ilu P, hipStream_t s) {
  const uint64_t v = (A->version << 32) ^ P->version;
  if (P->version32 == v && A->vcrs32.p && P->vcrs32.p) return;
  const size_t na = A->pat.ncrs * 9, nd = (size_t)A->nblk * 9, nl = P->ncrs * 9;
  A->vcrs32.alloc(na);
  A->vdia32.alloc(nd);
  P->vcrs32.alloc(nl);
  P->vdia32.alloc(nd);
  if (na)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(vec_grid(na)), dim3(256), 0, s,
                       A->vcrs32.p, A->vcrs.p, na);
  if (nd) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(vec_grid(nd)), dim3(256), 0, s,
                       A->vdia32.p, A->vdia.p, nd);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(vec_grid(nd)), dim3(256), 0, s,
                       P->vdia32.p, P->vdia.p, nd);
  }
  if (nl)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(vec_grid(nl)), dim3(256), 0, s,
                       P->vcrs32.p, P->vcrs.p, nl);
  check_launch("k_convert");
  P->version32 = v;
}

}  // namespace dfm2h

using namespace dfm2h;

static void need_solver(dfm2h_bsr A, dfm2h_ilu P) {
  DFM2H_REQUIRE(A && P, DFM2H_ERR_INVALID, "null handle");
  DFM2H_REQUIRE(A->has_pattern && P->has_pattern, DFM2H_ERR_STATE, "pattern not set");
  DFM2H_REQUIRE(A->nblk == P->nblk, DFM2H_ERR_INVALID, "matrix / preconditioner size mismatch");
}

extern "C" {

int dfm2h_vec_dot(const double *x, const double *y, size_t n, double *out) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    if (n == 0) {
      *out = 0.0;
      return;
    }
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    static thread_local DBuf<double> partials, res;
    static thread_local DBuf<unsigned> ticket;
    const unsigned g = vec_grid(n);
    if (partials.n < g) partials.alloc(kVecGridMax);
    if (!res.p) res.alloc(1);
    if (!ticket.p) {
      ticket.alloc(1);
      ticket.zero();
    }
    hipLaunchKernelGGL(k_dot, dim3(g), dim3(kVecBlock), 0, stream(), x, y, n, partials.p,
                       ticket.p, res.p);
    check_launch("k_dot");
    DFM2H_HIP(hipMemcpyAsync(out, res.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_vec_add_scaled(double *y, double alpha, const double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    hipLaunchKernelGGL(k_add_scaled, dim3(vec_grid(n)), dim3(256), 0, stream(), y, alpha, x, n);
    check_launch("k_add_scaled");
  });
}

int dfm2h_vec_scale_and_add(double *y, double beta, const double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x && y, DFM2H_ERR_INVALID, "null vector");
    hipLaunchKernelGGL(k_scale_and_add, dim3(vec_grid(n)), dim3(256), 0, stream(), y, beta, x, n);
    check_launch("k_scale_and_add");
  });
}

int dfm2h_vec_copy(double *dst, const double *src, size_t n) {
  return guard(__func__, [&] {
    if (!n || dst == src) return;
    DFM2H_REQUIRE(dst && src, DFM2H_ERR_INVALID, "null vector");
    DFM2H_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  });
}

int dfm2h_vec_set_zero(double *x, size_t n) {
  return guard(__func__, [&] {
    if (!n) return;
    DFM2H_REQUIRE(x, DFM2H_ERR_INVALID, "null vector");
    hipLaunchKernelGGL(k_fill<double>, dim3(vec_grid(n)), dim3(256), 0, stream(), x, 0.0, n);
    check_launch("k_fill");
  });
}

int dfm2h_solve_pcg(dfm2h_bsr A, dfm2h_ilu P, double *r, double *x, double tol,
                    unsigned max_nitr, double *hist, size_t hist_cap, size_t *n_hist) {
  return guard(__func__, [&] {
    need_solver(A, P);
    DFM2H_REQUIRE(r && x, DFM2H_ERR_INVALID, "null vector");
    PcgWork &W = work_of(P);
    PcgBufs<double> &B = W.f64;
    ensure_bufs(B, A, P, max_nitr);
    const size_t n = (size_t)A->nblk * 3;
    fence_in(W);
    DFM2H_HIP(hipMemcpyAsync(B.r.p, r, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
    const PcgArgs<double> a = make_args<double>(A, P, B, A->vcrs.p, A->vdia.p, P->vcrs.p, P->vdia.p);
    PcgScal st{};
    run_pcg(A, P, W, B, a, tol, max_nitr, st);
    DFM2H_HIP(hipMemcpyAsync(r, B.r.p, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
    DFM2H_HIP(hipMemcpyAsync(x, B.x.p, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
    // history: it+1 norms, plus the reference's final push when not converged
    std::vector<double> h((size_t)st.it + 1);
    DFM2H_HIP(hipMemcpyAsync(h.data(), B.hist.p, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, W.s));
    DFM2H_HIP(hipStreamSynchronize(W.s));
    if (!st.converged && !st.nan) h.push_back(h.back());
    if (hist)
      for (size_t k = 0; k < h.size() && k < hist_cap; ++k) hist[k] = h[k];
    if (n_hist) *n_hist = h.size();
    DFM2H_REQUIRE(!st.nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
  });
}

int dfm2h_solve_pcg_mixed(dfm2h_bsr A, dfm2h_ilu P, const double *b, double *x, double tol,
                          double inner_tol, unsigned max_outer, unsigned max_inner, double *hist,
                          size_t hist_cap, size_t *n_hist, size_t *n_inner) {
  return guard(__func__, [&] {
    need_solver(A, P);
    DFM2H_REQUIRE(b && x, DFM2H_ERR_INVALID, "null vector");
    PcgWork &W = work_of(P);
    PcgBufs<float> &B = W.f32;
    ensure_bufs(B, A, P, max_inner);
    const size_t n = (size_t)A->nblk * 3;
    const unsigned nbr = std::max(1u, ceil_div(A->nblk, kRowsPerBlock));
    if (W.mixed_partials.n < nbr) W.mixed_partials.alloc(nbr);
    if (!W.sq.p) W.sq.alloc(1);
    if (!W.ticket.p) {
      W.ticket.alloc(1);
      DFM2H_HIP(hipMemsetAsync(W.ticket.p, 0, sizeof(unsigned), stream()));
    }
    fence_in(W);
    ensure_f32(A, P, W.s);
    hipLaunchKernelGGL(k_fill<double>, dim3(vec_grid(n)), dim3(256), 0, W.s, x, 0.0, n);
    const PcgArgs<float> a =
        make_args<float>(A, P, B, A->vcrs32.p, A->vdia32.p, P->vcrs32.p, P->vdia32.p);
    std::vector<double> h;
    size_t inner_total = 0;
    bool nan = false;
    double sq0 = 0.0;
    for (unsigned k = 0; k <= max_outer; ++k) {
      // r = b - A x (FP64), rounded to FP32 for the inner solve
      hipLaunchKernelGGL(k_residual_to_f32, dim3(nbr), dim3(kBlock), 0, W.s, A->nblk,
                         A->pat.colind_d.p, A->pat.rowptr_d.p, A->vcrs.p, A->vdia.p, b, x,
                         B.r.p, W.mixed_partials.p, W.ticket.p, W.sq.p);
      check_launch("k_residual_to_f32");
      double sq = 0.0;
      DFM2H_HIP(hipMemcpyAsync(&sq, W.sq.p, sizeof(double), hipMemcpyDeviceToHost, W.s));
      DFM2H_HIP(hipStreamSynchronize(W.s));
      h.push_back(std::sqrt(sq));
      if (k == 0) {
        sq0 = sq;
        if (sq0 < 1.0e-30) break;
      }
      if (!(sq == sq) || sq > 1.0e308) {
        nan = true;
        break;
      }
      if (std::sqrt(sq / sq0) < tol || k == max_outer) break;
      PcgScal st{};
      run_pcg(A, P, W, B, a, inner_tol, max_inner, st);
      inner_total += (size_t)st.it;
      if (st.nan) {
        nan = true;
        break;
      }
      hipLaunchKernelGGL(k_add_f32_to_f64, dim3(vec_grid(n)), dim3(256), 0, W.s, x, B.x.p, n);
      check_launch("k_add_f32_to_f64");
    }
    DFM2H_HIP(hipEventRecord(W.ev_out, W.s));
    DFM2H_HIP(hipStreamWaitEvent(stream(), W.ev_out, 0));
    DFM2H_HIP(hipStreamSynchronize(W.s));
    if (hist)
      for (size_t k = 0; k < h.size() && k < hist_cap; ++k) hist[k] = h[k];
    if (n_hist) *n_hist = h.size();
    if (n_inner) *n_inner = inner_total;
    DFM2H_REQUIRE(!nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
  });
}

}  // extern "C"
