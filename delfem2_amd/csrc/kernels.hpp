// kernels.hpp -- gfx950 device code of the tet-FEM / BSR / ILU(0) path.
//
// Work mapping used by the block-row kernels (SpMV and both ILU sweeps):
// three lanes per 3x3 block row, one lane per output component, 21 rows per
// 64-lane wavefront (lane 63 idles).  Lane c of row i walks the row's blocks
// in stored order and reads row c of each block (24 contiguous bytes; the
// three lanes of a row cover the block's 72 bytes), so the per-component sum
// keeps the reference's association and order exactly.
#pragma once

#include "common.hpp"

namespace dfm2h {

constexpr int kRowsPerWave = 21;
constexpr int kBlock = 256;  // threads per workgroup for the row kernels
constexpr int kRowsPerBlock = kRowsPerWave * (kBlock / kWave);  // 84

struct RowLane {
  unsigned r;  // row slot inside the wave (0..21; 21 = idle lane)
  unsigned c;  // component 0..2
  unsigned slot;  // row slot inside the grid
};

__device__ inline RowLane row_lane() {
  const unsigned lane = threadIdx.x & (kWave - 1);
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  RowLane L;
  L.r = lane / 3;
  L.c = lane - 3 * L.r;
  L.slot = wave * kRowsPerWave + L.r;
  return L;
}

// one component of a 3x3 block times a 3-vector: (v0*x0 + v1*x1) + v2*x2
template <typename T>
__device__ inline T blk_row_dot(const T *__restrict__ v, const T *__restrict__ x) {
  return v[0] * x[0] + v[1] * x[1] + v[2] * x[2];
}

// ---------------------------------------------------------------------------
// y = alpha*A*x + beta*y (ls_block_sparse.cpp:215-273, :103-133)

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_bsr_matvec(unsigned nblk, const unsigned *__restrict__ colind,
             const unsigned *__restrict__ rowptr, const T *__restrict__ vcrs,
             const T *__restrict__ vdia, T *__restrict__ y, T alpha,
             const T *__restrict__ x, T beta) {
  const RowLane L = row_lane();
  if (L.r >= kRowsPerWave || L.slot >= nblk) return;
  const unsigned i = L.slot, c = L.c;
  T acc = y[3 * (size_t)i + c] * beta;
  const unsigned s1 = colind[i + 1];
  for (unsigned s = colind[i]; s < s1; ++s) {
    const size_t j = rowptr[s];
    acc += alpha * blk_row_dot(vcrs + (size_t)s * 9 + 3 * c, x + 3 * j);
  }
  acc += alpha * blk_row_dot(vdia + (size_t)i * 9 + 3 * c, x + 3 * (size_t)i);
  y[3 * (size_t)i + c] = acc;
}

// ---------------------------------------------------------------------------
// ILU(0) forward sweep of one dependency level (ls_ilu_block_sparse.cpp:544-573):
//   t_i = src_i - sum_{j<i} L_ij vec_j  (stored order),  vec_i = D_i^-1 t_i
// Rows of one level never read each other, so they run in parallel.  src may
// equal vec (in place, as the reference) or hold the right-hand side while
// vec collects the result (fuses the reference's copy Pr = r).

template <typename T>
__device__ inline void fwd_row(unsigned i, unsigned c, bool valid, const unsigned *__restrict__ colind,
                               const unsigned *__restrict__ rowptr,
                               const unsigned *__restrict__ diaind, const T *__restrict__ vcrs,
                               const T *__restrict__ vdia, const T *src, T *vec) {
  T t = 0;
  if (valid) {
    t = src[3 * (size_t)i + c];
    const unsigned s1 = diaind[i];
    for (unsigned s = colind[i]; s < s1; ++s) {
      const size_t j = rowptr[s];
      t -= blk_row_dot(vcrs + (size_t)s * 9 + 3 * c, vec + 3 * j);
    }
  }
  // gather the row's three components from its three lanes
  const int lane = threadIdx.x & (kWave - 1);
  const int base = lane - (int)c;
  const T t0 = __shfl(t, base + 0, kWave);
  const T t1 = __shfl(t, base + 1, kWave);
  const T t2 = __shfl(t, base + 2, kWave);
  if (valid) {
    const T *d = vdia + (size_t)i * 9 + 3 * c;
    vec[3 * (size_t)i + c] = d[0] * t0 + d[1] * t1 + d[2] * t2;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ilu_fwd_level(const unsigned *__restrict__ rows, unsigned nrows,
                const unsigned *__restrict__ colind, const unsigned *__restrict__ rowptr,
                const unsigned *__restrict__ diaind, const T *__restrict__ vcrs,
                const T *__restrict__ vdia, const T *src, T *vec) {
  const RowLane L = row_lane();
  const bool valid = L.r < kRowsPerWave && L.slot < nrows;
  const unsigned i = valid ? rows[L.slot] : 0u;
  fwd_row(i, L.c, valid, colind, rowptr, diaind, vcrs, vdia, src, vec);
}

// ILU(0) backward sweep of one level (ls_ilu_block_sparse.cpp:721-749):
//   vec_i = vec_i - sum_{j>i} U_ij vec_j   (U already scaled by D^-1)
template <typename T>
__device__ inline T bwd_row(unsigned i, unsigned c, const unsigned *__restrict__ colind,
                            const unsigned *__restrict__ rowptr,
                            const unsigned *__restrict__ diaind, const T *__restrict__ vcrs,
                            T *vec) {
  T t = vec[3 * (size_t)i + c];
  const unsigned s1 = colind[i + 1];
  for (unsigned s = diaind[i]; s < s1; ++s) {
    const size_t j = rowptr[s];
    t -= blk_row_dot(vcrs + (size_t)s * 9 + 3 * c, vec + 3 * j);
  }
  vec[3 * (size_t)i + c] = t;
  return t;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ilu_bwd_level(const unsigned *__restrict__ rows, unsigned nrows,
                const unsigned *__restrict__ colind, const unsigned *__restrict__ rowptr,
                const unsigned *__restrict__ diaind, const T *__restrict__ vcrs, T *vec) {
  const RowLane L = row_lane();
  if (L.r >= kRowsPerWave || L.slot >= nrows) return;
  bwd_row(rows[L.slot], L.c, colind, rowptr, diaind, vcrs, vec);
}

}  // namespace dfm2h
