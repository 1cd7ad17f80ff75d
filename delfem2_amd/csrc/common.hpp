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

hipStream_t stream();
// checks the last kernel launch
void check_launch(const char *what);

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
// handles

struct Pattern {
  unsigned nblk = 0;
  size_t ncrs = 0;
  std::vector<unsigned> colind, rowptr;  // host copies
  DBuf<unsigned> colind_d, rowptr_d;
  DBuf<unsigned> rowof_d;  // block row of each crs slot
};

}  // namespace dfm2h

struct dfm2h_bsr_s {
  unsigned nblk = 0;
  bool has_pattern = false;
  dfm2h::Pattern pat;
  dfm2h::DBuf<double> vcrs, vdia;
  // FP32 shadow of the values for mixed precision (made on demand)
  dfm2h::DBuf<float> vcrs32, vdia32;
  uint64_t version = 0;  // bumped whenever values change
};

struct dfm2h_tetmesh_s {
  size_t ntet = 0, np = 0;
  dfm2h::DBuf<unsigned> tet;        // [4*ntet]
  dfm2h::DBuf<unsigned> elsup_ind;  // [np+1]
  dfm2h::DBuf<unsigned> elsup;      // element ids, ascending per node
};

namespace dfm2h {
struct PcgWork;  // pcg.hip
void destroy_pcg_work(PcgWork *);
}  // namespace dfm2h

struct dfm2h_ilu_s {
  unsigned nblk = 0;
  size_t ncrs = 0;
  bool has_pattern = false;
  bool decomposed = false;
  std::vector<unsigned> colind, rowptr, diaind;  // host (sorted pattern)
  dfm2h::DBuf<unsigned> colind_d, rowptr_d, diaind_d;
  dfm2h::DBuf<double> vcrs, vdia;
  // A slot -> ILU slot map for CopyValue (built against `mapped_for`)
  const dfm2h_bsr_s *mapped_for = nullptr;
  dfm2h::DBuf<int> a2l_d;
  // level schedules: rows grouped by dependency level, ascending inside a level
  std::vector<unsigned> levptr_f, levptr_b;
  dfm2h::DBuf<unsigned> levrows_f, levrows_b;
  dfm2h::DBuf<int> nsing_d;
  // FP32 copies of the factors (mixed precision)
  dfm2h::DBuf<float> vcrs32, vdia32;
  uint64_t version = 0;
  uint64_t version32 = ~0ull;
  dfm2h::PcgWork *work = nullptr;
  ~dfm2h_ilu_s() { dfm2h::destroy_pcg_work(work); }
};

namespace dfm2h {

// ---------------------------------------------------------------------------
// device helpers

constexpr int kWave = 64;

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

// Last-block-done detection for single-pass grid reductions (the
// threadfence-reduction pattern).  Every block calls it after writing its
// partial; exactly one block gets true, after all partials are visible.
__device__ inline bool last_block(unsigned *ticket, unsigned nblocks) {
  __shared__ bool am_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = atomicAdd(ticket, 1u);
    am_last = (t == nblocks - 1);
  }
  __syncthreads();
  if (am_last) __threadfence();
  return am_last;
}

// Deterministic sum of partials[0..n) by one block of BS threads (strided
// per-thread sums, then a fixed tree).
template <int BS>
__device__ inline double reduce_partials(const double *partials, unsigned n) {
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += BS) s += partials[i];
  return block_reduce_sum<BS>(s);
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace dfm2h
