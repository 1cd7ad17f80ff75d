// common.hpp -- shared state, error plumbing and device helpers of the
// gfx950 tet-FEM / BSR / ILU(0) / PCG library.
//
// All device code is built with -ffp-contract=off: no multiply-add is fused,
// so every product and sum rounds exactly as in the (x86-64, non-FMA)
// reference and the FP64 assembly, MatVec, factorisation and triangular
// sweeps are bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dfm2_hip.h"

namespace dfm2h {

// ---------------------------------------------------------------------------
// errors

void set_error(const std::string &msg);

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define DFM2H_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::dfm2h::Error(DFM2H_ERR_HIP, std::string(#call) + ": " +                \
                                              hipGetErrorString(e_));                \
  } while (0)

#define DFM2H_REQUIRE(cond, code, msg)                                \
  do {                                                                \
    if (!(cond)) throw ::dfm2h::Error((code), std::string(msg));      \
  } while (0)

// Runs f() and maps exceptions to status codes; never lets one escape.
template <class F>
int guard(const char *fn, F &&f) {
  try {
    f();
    return DFM2H_OK;
  } catch (const Error &e) {
    set_error(std::string(fn) + ": " + e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error(std::string(fn) + ": host allocation failed");
    return DFM2H_ERR_INVALID;
  } catch (const std::exception &e) {
    set_error(std::string(fn) + ": " + e.what());
    return DFM2H_ERR_INVALID;
  } catch (...) {
    set_error(std::string(fn) + ": unknown error");
    return DFM2H_ERR_INVALID;
  }
}

// the calling thread's stream (dfm2h_set_stream)
hipStream_t stream();
// process-wide unique serial numbers (meshes, patterns): a handle created
// at the address of a destroyed one never carries the same serial
uint64_t next_serial();
// Throws DFM2H_ERR_INVALID unless p (when non-null) is memory a kernel can
// read: device, managed, or pinned host memory.  Pageable host memory --
// e.g. a std::vector's data() passed where the reference took a host
// array -- is rejected before anything is launched.
void require_device(const void *p, const char *what);
#define DFM2H_DEVICE_ARG(p) ::dfm2h::require_device((p), #p)
// checks the last kernel launch
void check_launch(const char *what);

// Autotuner cache: (solver kind, floor(log2 ndof), fp32) -> launch variant.
// Kept for the process; when the environment variable DFM2H_TUNE_CACHE
// names a file, entries are read from it once and new ones appended
// ("kind bucket fp32 variant ms" per line).  Lookup returns -1 when absent
// or when the stored variant is not below n_variants.
int tune_lookup(const char *kind, int bucket, int fp32, int n_variants);
void tune_store(const char *kind, int bucket, int fp32, int variant, double ms);
int dof_bucket(size_t ndof);

// ---------------------------------------------------------------------------
// timing events (RAII over hipEventCreate; destroyed on every exit path)

template <int N>
struct TimingEvents {
  hipEvent_t e[N] = {};
  TimingEvents() {
    for (auto &x : e) {
      const hipError_t err = hipEventCreate(&x);
      if (err != hipSuccess) {
        x = nullptr;
        destroy();
        DFM2H_HIP(err);
      }
    }
  }
  TimingEvents(const TimingEvents &) = delete;
  TimingEvents &operator=(const TimingEvents &) = delete;
  ~TimingEvents() { destroy(); }
  void destroy() {
    for (auto &x : e)
      if (x) (void)hipEventDestroy(x);
    for (auto &x : e) x = nullptr;
  }
  hipEvent_t operator[](int i) const { return e[i]; }
};

// ---------------------------------------------------------------------------
// device buffer (RAII over hipMalloc)

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    DFM2H_HIP(hipMalloc(&p, sizeof(T) * count));
    n = count;
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) DFM2H_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, stream()));
  }
  void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
  void download(T *h, size_t count) const {
    if (count) DFM2H_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, stream()));
  }
  void zero() {
    if (n) DFM2H_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, stream()));
  }
};



// ---------------------------------------------------------------------------
// device helpers

constexpr int kWave = 64;

// Inner iterations after the last reliable update that halved the smallest
// FP64 residual so far after which a mixed solve counts as stalled and is
// finished by the FP64 solver (checked at every reliable update; the forced
// updates check it at least every period).
constexpr int kStallIters = 500;

__device__ inline double wave_reduce_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Fixed-shape block reduction (deterministic for a fixed block size).
// The result is valid in every thread.
template <int BS>
__device__ inline double block_reduce_sum(double v) {
  static_assert(BS % kWave == 0, "block size must be a multiple of the wave size");
  constexpr int NW = BS / kWave;
  __shared__ double sm[NW];
  v = wave_reduce_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) t += sm[w];
  __syncthreads();
  return t;
}

// Deterministic sum of partials[0..n) by one block of BS threads (strided
// per-thread sums, then a fixed tree).  Grid reductions write one partial per
// workgroup and a separate one-workgroup kernel sums them: the kernel
// boundary makes the partials visible without a device-scope fence in every
// workgroup (which on a multi-XCD part writes back L2 each time: measured at
// 823,875 DOF, the PCG's alpha step run by the SpMV's last workgroup, with a
// release fence and a counter increment in each of its 3,270 workgroups,
// took the mixed solve from 58.9 to 104.8 ms and the FP64 solve from 104.0
// to 180.6 ms, same bits).
template <int BS>
__device__ inline double reduce_partials(const double *partials, unsigned n) {
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += BS) s += partials[i];
  return block_reduce_sum<BS>(s);
}

// The same sums for NS partial sets at once (set k: p[k][0..n[k]), n[k] = 0
// for an unused set), each formed in exactly reduce_partials' order, with
// the loads of U strides of every set issued before their adds: the
// partials were just written by other XCDs, so every load round is a full
// memory latency, and one round replaces up to U x NS of them.  Neither U
// nor the other sets change a set's sum.  partial_sums_thread forms each
// thread's share, partial_sums_block the block's; reduce_partial_sets both.
template <int BS, int NS, int U = 8>
__device__ inline void partial_sums_thread(const double *const (&p)[NS], const unsigned (&n)[NS],
                                           double (&s)[NS]) {
  unsigned nmax = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    s[k] = 0.0;
    nmax = n[k] > nmax ? n[k] : nmax;
  }
  for (unsigned base = 0; base < nmax; base += BS * U) {
    double v[NS][U];
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const unsigned i = base + u * BS + threadIdx.x;
        v[k][u] = i < n[k] ? p[k][i] : 0.0;
      }
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (base + u * BS + threadIdx.x < n[k]) s[k] += v[k][u];
  }
}

// The block's sums of the per-thread sums s (fixed wave, then fixed wave
// order), in every thread.
template <int BS, int NS>
__device__ inline void partial_sums_block(const double (&s)[NS], double (&t)[NS]) {
  constexpr int NW = BS / kWave;
  __shared__ double sm[NS][NW];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double w = wave_reduce_sum(s[k]);
    if (lane == 0) sm[k][wid] = w;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) a += sm[k][w];
    t[k] = a;
  }
  __syncthreads();
}

template <int BS, int NS, int U = 8>
__device__ inline void reduce_partial_sets(const double *const (&p)[NS], const unsigned (&n)[NS],
                                           double (&t)[NS]) {
  double s[NS];
  partial_sums_thread<BS, NS, U>(p, n, s);
  partial_sums_block<BS, NS>(s, t);
}

// Solver state written by earlier kernels (flags, step lengths) is read with
// vector loads at agent scope, never through the scalar data cache: a
// uniform load of it can otherwise return a scalar-cache line filled before
// the write (measured: the mixed PCG's one-workgroup steps, loading the
// whole state that way, saw stale restart / rmax2 fields and took other
// iterates).
template <typename V>
__device__ __attribute__((always_inline)) inline V ld_state(const V *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace dfm2h
