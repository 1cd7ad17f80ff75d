// krylov.hip -- ILU(0)-preconditioned BiCGStab (the method of
// vecxitrsol.h:350-437 Solve_PBiCGStab) for the non-symmetric systems of
// the stiffness-warped time step, run on the device: the iteration is a
// fixed chain of kernels captured in HIP graphs of kChunk iterations, every
// scalar (the step lengths, the stop test) lives in HBM, and the host only
// polls a done flag, one graph behind.
//
// One iteration, with M the ILU(0) factors, A the matrix, r^ the shadow
// residual (the start residual, kept), in this order of operations:
//   z1 = M^-1 p                 forward + backward sweeps (the reference's
//                               summation order)
//   q  = A z1                   MatVec (y scaled by 0 first, stored block
//                               order, diagonal last); (q, r^) per workgroup
//   alpha = rho / (q, r^)       rho = (r, r^)
//   s  = r; s += -alpha q
//   z2 = M^-1 s
//   t  = A z2                   (t, t) and (s, t) per workgroup
//   omega = (s, t) / (t, t)
//   x += alpha z1; x += omega z2; r = s; r += -omega t
//                               (r, r) and (r, r^) per workgroup
//   history |r| / |r_0|; stop below the tolerance (or NaN, or the cap);
//   beta = (r, r^) alpha / (rho omega); rho = (r, r^)
//   p *= beta; p += r; p += (-beta omega) q
// Each vector update rounds exactly as the reference's elementwise loops;
// the dot products are sums in a fixed tree order (deterministic, not the
// reference's sequential order), so iterates agree with the reference to
// rounding.
#include "handles.hpp"

#include <cmath>

namespace dfm2h {

constexpr int kKryBlock = 256;
constexpr unsigned kKryGridMax = 2048;
// iterations per graph launch: the host polls one graph behind, so up to
// two chunks run past convergence as immediate returns (24 launches each):
// 2 keeps that below 0.1 ms, and a chunk of 2 iterations (~0.4 ms at
// 823,875 DOF) still hides the host's launch and poll
constexpr unsigned kChunk = 2;
constexpr int kScalThreads = 1024;

static unsigned kry_grid(size_t n) {
  return std::max(1u, std::min(kKryGridMax, ceil_div(n, kKryBlock)));
}

struct BicgScal {
  double rho;  // (r, r^) of the current residual
  double alpha, omega, beta;
  double inv_sq0, tol;
  int it, done, converged, nan, max_it, pad;
};

struct BicgArgs {
  size_t n;
  // vectors: residual, solution, shadow residual, direction, the two
  // preconditioned vectors, their products with A, the intermediate residual
  double *r, *x, *rs, *p, *z1, *q, *z2, *t, *s;
  double *part_a, *part_b;  // two sets of per-workgroup partial sums
  double *hist;
  BicgScal *S;
  // the matrix in the sliced layout
  PlanView plan;
  unsigned nsl;
  const double *val, *dia;
};

// x = 0, r^ = r, p = r, partial (r, r)
__global__ void __launch_bounds__(kKryBlock) k_bicg_init(BicgArgs a) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * kKryBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kKryBlock) {
    const double v = a.r[i];
    a.x[i] = 0.0;
    a.rs[i] = v;
    a.p[i] = v;
    s += v * v;
  }
  s = block_reduce_sum<kKryBlock>(s);
  if (threadIdx.x == 0) a.part_a[blockIdx.x] = s;
}

enum BicgStep { kBicgStart = 0, kBicgAlpha = 1, kBicgOmega = 2, kBicgNext = 3 };

// One workgroup: the scalar step after a reduction.  na / nb partials in
// part_a / part_b.
template <int STEP>
__global__ void __launch_bounds__(kScalThreads)
k_bicg_scalar(BicgScal *S, const double *__restrict__ pa, unsigned na,
              const double *__restrict__ pb, unsigned nb, double *hist) {
  if (STEP != kBicgStart && ld_state(&S->done)) return;
  const double *const ps[2] = {pa, pb};
  const unsigned ns[2] = {na, STEP == kBicgOmega || STEP == kBicgNext ? nb : 0u};
  double t[2];
  reduce_partial_sets<kScalThreads, 2>(ps, ns, t);
  if (threadIdx.x != 0) return;
  if (STEP == kBicgStart) {  // t[0] = |r0|^2
    S->it = 0;
    S->converged = 0;
    S->nan = 0;
    S->done = 0;
    if (t[0] < 1.0e-60) {  // the reference's early return: one history entry
      hist[0] = sqrt(t[0]);
      S->done = 1;
      S->converged = 1;
    } else if (!(t[0] == t[0]) || t[0] > 1.0e308) {
      S->done = 1;
      S->nan = 1;
    } else if (ld_state(&S->max_it) <= 0) {
      S->done = 1;
    }
    S->inv_sq0 = 1.0 / t[0];
    S->rho = t[0];  // (r, r^) with r^ = r
  } else if (STEP == kBicgAlpha) {  // t[0] = (q, r^)
    S->alpha = ld_state(&S->rho) / t[0];
  } else if (STEP == kBicgOmega) {  // t[0] = (t, t), t[1] = (s, t)
    S->omega = t[1] / t[0];
  } else {  // t[0] = (r, r), t[1] = (r, r^)
    const int it = ld_state(&S->it) + 1;
    S->it = it;
    const double ratio = sqrt(t[0] * ld_state(&S->inv_sq0));
    hist[it - 1] = ratio;
    if (ratio < ld_state(&S->tol)) {
      S->done = 1;
      S->converged = 1;
    } else if (!(ratio == ratio) || ratio > 1.0e308) {
      S->done = 1;
      S->nan = 1;
    } else if (it >= ld_state(&S->max_it)) {
      S->done = 1;
    } else {
      S->beta = t[1] * ld_state(&S->alpha) / (ld_state(&S->rho) * ld_state(&S->omega));
      S->rho = t[1];
    }
  }
}

// y = A z (y scaled by 0 first, then the stored blocks in order, then the
// diagonal: k_sliced_matvec's arithmetic) and per-workgroup partials of
// (y, u) into pa and, with TWO, of (y, y) into pa and (w, y) into pb.
template <bool TWO>
__global__ void __launch_bounds__(kPackBlock)
k_bicg_matvec(BicgArgs a, const double *z, double *y, const double *u, const double *w) {
  if (ld_state(&a.S->done)) return;
  const SliceLane L = slice_lane(0);
  double s0 = 0.0, s1 = 0.0;
  if (L.s < a.nsl && L.c < 3) {
    const unsigned i = a.plan.row[L.s * kSliceRows + L.r];
    if (i != ~0u) {
      const size_t e = 3 * (size_t)i + L.c;
      double acc = y[e] * 0.0;
      acc = sliced_row_sum<double, false, unroll_spmv<double>(), 3>(
          acc, 1.0, L, a.plan.len[L.s * kSliceRows + L.r], a.plan.kmax[L.s], a.plan.off[L.s],
          a.plan.col, a.val, z);
      const Trip<double> zi = load_node<double, 3>(z, i);
      acc += 1.0 * dia_row_dot(a.dia, L, zi.a, zi.b, zi.c);
      y[e] = acc;
      if (TWO) {
        s0 = acc * acc;
        s1 = w[e] * acc;
      } else {
        s0 = acc * u[e];
      }
    }
  }
  s0 = block_reduce_sum<kPackBlock>(s0);
  if (TWO) s1 = block_reduce_sum<kPackBlock>(s1);
  if (threadIdx.x == 0) {
    a.part_a[blockIdx.x] = s0;
    if (TWO) a.part_b[blockIdx.x] = s1;
  }
}

// s = r; s += -alpha q
__global__ void __launch_bounds__(kKryBlock) k_bicg_s(BicgArgs a) {
  if (ld_state(&a.S->done)) return;
  const double nalpha = -ld_state(&a.S->alpha);
  for (size_t i = (size_t)blockIdx.x * kKryBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kKryBlock)
    a.s[i] = a.r[i] + nalpha * a.q[i];
}

// x += alpha z1; x += omega z2; r = s; r += -omega t; partials (r, r) and
// (r, r^)
__global__ void __launch_bounds__(kKryBlock) k_bicg_xr(BicgArgs a) {
  if (ld_state(&a.S->done)) return;
  const double alpha = ld_state(&a.S->alpha), omega = ld_state(&a.S->omega);
  double srr = 0.0, srs = 0.0;
  for (size_t i = (size_t)blockIdx.x * kKryBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kKryBlock) {
    double xi = a.x[i] + alpha * a.z1[i];
    xi = xi + omega * a.z2[i];
    a.x[i] = xi;
    const double ri = a.s[i] + (-omega) * a.t[i];
    a.r[i] = ri;
    srr += ri * ri;
    srs += ri * a.rs[i];
  }
  srr = block_reduce_sum<kKryBlock>(srr);
  srs = block_reduce_sum<kKryBlock>(srs);
  if (threadIdx.x == 0) {
    a.part_a[blockIdx.x] = srr;
    a.part_b[blockIdx.x] = srs;
  }
}

// p *= beta; p += 1.0 r; p += (-beta omega) q
__global__ void __launch_bounds__(kKryBlock) k_bicg_p(BicgArgs a) {
  if (ld_state(&a.S->done)) return;
  const double beta = ld_state(&a.S->beta), bw = -beta * ld_state(&a.S->omega);
  for (size_t i = (size_t)blockIdx.x * kKryBlock + threadIdx.x; i < a.n;
       i += (size_t)gridDim.x * kKryBlock) {
    double pi = a.p[i] * beta;
    pi = pi + 1.0 * a.r[i];
    a.p[i] = pi + bw * a.q[i];
  }
}

// Buffers, stream and captured graphs of one preconditioner's BiCGStab.
struct BicgWork {
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_poll[2] = {nullptr, nullptr};
  BicgScal *host_status = nullptr;  // pinned, 2 entries
  DBuf<double> r, x, rs, p, z1, q, z2, t, sv, part_a, part_b, hist;
  DBuf<BicgScal> S;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> sig;
  BicgWork() {
    DFM2H_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_poll[0], hipEventDisableTiming));
    DFM2H_HIP(hipEventCreateWithFlags(&ev_poll[1], hipEventDisableTiming));
    DFM2H_HIP(hipHostMalloc((void **)&host_status, 2 * sizeof(BicgScal), hipHostMallocDefault));
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    sig.clear();
  }
  ~BicgWork() {
    drop_graph();
    if (host_status) (void)hipHostFree(host_status);
    for (hipEvent_t e : {ev_in, ev_poll[0], ev_poll[1]})
      if (e) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
};

void destroy_bicg_work(BicgWork *w) { delete w; }

// One iteration of the chain above on stream st.
static void launch_bicg_iteration(dfm2h_ilu P, const BicgArgs &a, hipStream_t st) {
  const unsigned gv = kry_grid(a.n), gs = slice_grid(a.nsl);
  int *done = &a.S->done;
  ilu_forward_levels_double(P, a.p, a.z1, st, done);
  ilu_backward_levels_double(P, a.z1, st, done);
  hipLaunchKernelGGL(k_bicg_matvec<false>, dim3(gs), dim3(kPackBlock), 0, st, a, a.z1, a.q, a.rs,
                     (const double *)nullptr);
  hipLaunchKernelGGL(k_bicg_scalar<kBicgAlpha>, dim3(1), dim3(kScalThreads), 0, st, a.S,
                     a.part_a, gs, (const double *)nullptr, 0u, a.hist);
  hipLaunchKernelGGL(k_bicg_s, dim3(gv), dim3(kKryBlock), 0, st, a);
  ilu_forward_levels_double(P, a.s, a.z2, st, done);
  ilu_backward_levels_double(P, a.z2, st, done);
  hipLaunchKernelGGL(k_bicg_matvec<true>, dim3(gs), dim3(kPackBlock), 0, st, a, a.z2, a.t,
                     (const double *)nullptr, a.s);
  hipLaunchKernelGGL(k_bicg_scalar<kBicgOmega>, dim3(1), dim3(kScalThreads), 0, st, a.S,
                     a.part_a, gs, a.part_b, gs, a.hist);
  hipLaunchKernelGGL(k_bicg_xr, dim3(gv), dim3(kKryBlock), 0, st, a);
  hipLaunchKernelGGL(k_bicg_scalar<kBicgNext>, dim3(1), dim3(kScalThreads), 0, st, a.S,
                     a.part_a, gv, a.part_b, gv, a.hist);
  hipLaunchKernelGGL(k_bicg_p, dim3(gv), dim3(kKryBlock), 0, st, a);
}

}  // namespace dfm2h

using namespace dfm2h;

extern "C" {

int dfm2h_solve_pbicgstab(dfm2h_bsr A, dfm2h_ilu P, double *r, double *x, double tol,
                          unsigned max_niter, double *hist, size_t hist_cap, size_t *n_hist) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A && P && r && x && n_hist, DFM2H_ERR_INVALID, "null argument");
    DFM2H_REQUIRE(hist || hist_cap == 0, DFM2H_ERR_INVALID, "null history with a capacity");
    DFM2H_REQUIRE(A->has_pattern && P->has_pattern, DFM2H_ERR_STATE, "pattern not set");
    DFM2H_REQUIRE(A->nblk == P->nblk, DFM2H_ERR_INVALID, "matrix / preconditioner size mismatch");
    DFM2H_DEVICE_ARG(r);
    DFM2H_DEVICE_ARG(x);
    flush_merge(A);
    if (!P->bicg) P->bicg = new BicgWork;
    BicgWork &W = *P->bicg;
    const size_t n = 3 * (size_t)A->nblk;
    // at most max_niter - 1 iterations, one history entry each (the
    // reference's loop runs iitr = 1 .. max_niter - 1)
    const unsigned max_it = max_niter > 0 ? max_niter - 1 : 0;
    const unsigned gv = kry_grid(n), gs = slice_grid(A->plan.nslices);
    if (W.r.n != n) {
      for (DBuf<double> *v : {&W.r, &W.x, &W.rs, &W.p, &W.z1, &W.q, &W.z2, &W.t, &W.sv}) {
        v->alloc(n);
        v->zero();  // q and t start at zero, as the reference's vectors
      }
    }
    if (W.part_a.n < std::max(gv, gs)) {
      W.part_a.alloc(std::max(gv, gs));
      W.part_b.alloc(std::max(gv, gs));
    }
    if (W.hist.n < (size_t)max_it + 1) W.hist.alloc((size_t)max_it + 1);
    if (!W.S.p) W.S.alloc(1);
    // the caller's residual in, on the solver stream after the caller's work
    DFM2H_HIP(hipEventRecord(W.ev_in, stream()));
    DFM2H_HIP(hipStreamWaitEvent(W.s, W.ev_in, 0));
    if (n)
      DFM2H_HIP(hipMemcpyAsync(W.r.p, r, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
    ensure_packed(A->plan, A->pk64, A->vcrs.p, A->vdia.p, A->version, W.s);
    BicgScal init{};
    init.tol = tol;
    init.max_it = (int)max_it;
    DFM2H_HIP(hipMemcpyAsync(W.S.p, &init, sizeof(BicgScal), hipMemcpyHostToDevice, W.s));
    BicgArgs a{};
    a.n = n;
    a.r = W.r.p;
    a.x = W.x.p;
    a.rs = W.rs.p;
    a.p = W.p.p;
    a.z1 = W.z1.p;
    a.q = W.q.p;
    a.z2 = W.z2.p;
    a.t = W.t.p;
    a.s = W.sv.p;
    a.part_a = W.part_a.p;
    a.part_b = W.part_b.p;
    a.hist = W.hist.p;
    a.S = W.S.p;
    a.plan = view(A->plan);
    a.nsl = A->plan.nslices;
    a.val = A->pk64.val.p;
    a.dia = A->pk64.dia.p;
    hipLaunchKernelGGL(k_bicg_init, dim3(gv), dim3(kKryBlock), 0, W.s, a);
    hipLaunchKernelGGL(k_bicg_scalar<kBicgStart>, dim3(1), dim3(kScalThreads), 0, W.s, W.S.p,
                       W.part_a.p, gv, (const double *)nullptr, 0u, W.hist.p);
    check_launch("bicgstab start");
    // the factors' packed copies (before the capture, which must not pack),
    // then the graph of kChunk iterations, rebuilt when a buffer or a
    // pattern changed
    ensure_packed(P->fplan, P->f64, P->vcrs.p, P->vdia.p, P->version, W.s);
    ensure_packed(P->bplan, P->b64, P->vcrs.p, nullptr, P->version, W.s);
    const std::vector<uintptr_t> sig = {
        (uintptr_t)A,           (uintptr_t)A->pattern_serial, (uintptr_t)P->pattern_serial,
        (uintptr_t)a.val,       (uintptr_t)a.dia,             (uintptr_t)P->f64.val.p,
        (uintptr_t)P->f64.dia.p, (uintptr_t)P->b64.val.p,     (uintptr_t)W.r.p,
        (uintptr_t)W.part_a.p,  (uintptr_t)W.hist.p,          (uintptr_t)W.S.p};
    if (!W.exec || W.sig != sig) {
      W.drop_graph();
      DFM2H_HIP(hipStreamBeginCapture(W.s, hipStreamCaptureModeThreadLocal));
      for (unsigned k = 0; k < kChunk; ++k) launch_bicg_iteration(P, a, W.s);
      const hipError_t e = hipStreamEndCapture(W.s, &W.graph);
      DFM2H_REQUIRE(e == hipSuccess, DFM2H_ERR_HIP,
                    std::string("graph capture: ") + hipGetErrorString(e));
      DFM2H_HIP(hipGraphInstantiate(&W.exec, W.graph, nullptr, nullptr, 0));
      W.sig = sig;
    }
    // graphs until the device says done, polled one graph behind
    bool finished = false;
    unsigned launched = 0;
    for (unsigned c = 0; launched < max_it && !finished; ++c) {
      DFM2H_HIP(hipGraphLaunch(W.exec, W.s));
      launched += kChunk;
      DFM2H_HIP(hipMemcpyAsync(&W.host_status[c & 1], W.S.p, sizeof(BicgScal),
                               hipMemcpyDeviceToHost, W.s));
      DFM2H_HIP(hipEventRecord(W.ev_poll[c & 1], W.s));
      if (c == 0) continue;
      DFM2H_HIP(hipEventSynchronize(W.ev_poll[(c - 1) & 1]));
      finished = W.host_status[(c - 1) & 1].done != 0;
    }
    BicgScal st{};
    DFM2H_HIP(hipMemcpyAsync(&st, W.S.p, sizeof(BicgScal), hipMemcpyDeviceToHost, W.s));
    if (n) {
      DFM2H_HIP(hipMemcpyAsync(r, W.r.p, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
      DFM2H_HIP(hipMemcpyAsync(x, W.x.p, n * sizeof(double), hipMemcpyDeviceToDevice, W.s));
    }
    DFM2H_HIP(hipStreamSynchronize(W.s));
    // history: one entry per iteration, or the single |r0| of an early return
    const size_t nh = st.it == 0 && st.converged ? 1 : (size_t)st.it;
    std::vector<double> h(nh);
    if (nh) {
      DFM2H_HIP(hipMemcpyAsync(h.data(), W.hist.p, nh * sizeof(double), hipMemcpyDeviceToHost,
                               W.s));
      DFM2H_HIP(hipStreamSynchronize(W.s));
    }
    for (size_t k = 0; k < nh && k < hist_cap; ++k) hist[k] = h[k];
    *n_hist = nh;
    DFM2H_REQUIRE(!st.nan, DFM2H_ERR_NAN, "residual became NaN or Inf");
  });
}

}  // extern "C"
