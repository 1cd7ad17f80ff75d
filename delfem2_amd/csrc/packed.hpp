// packed.hpp -- sliced, wavefront-coalesced storage of a set of 3x3 block
// rows, and the row kernels that run on it (SpMV, ILU forward / backward).
//
// A "plan" groups rows (all rows for the SpMV; the rows of one dependency
// level for a sweep) into slices of 21 rows.  One wavefront owns a slice:
// lane = c*21 + r handles component c of slice row r (lane 63 idles).  For
// the k-th block of every row of slice s (block-slot q = off[s] + k):
//   col [q*21 + r]                      column of row r's k-th block
//   val                                 entry (c, t) of that block, for
//                                       lane = c*21 + r at
//     FP64: [(q*63 + lane)*3 + t]       (a lane's block row is contiguous)
//     FP32: [q*189 + t*63 + lane]       (planar)
// so every value load of a wavefront covers one contiguous run, and the
// column index is shared by the three lanes of a row.  The row's own 3x3
// block (diagonal of A, or D^-1 of the factorisation) is held the same way
// with s in place of q.  Vectors are read with a node stride VS: 3 (the
// reference's AoS layout) or 4 (padded).
// Blocks of a row keep their CRS order, so every per-component sum is formed
// in exactly the order of the reference loops (bit-identical in FP64).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace dfm2h {

// Value layout inside a block-slot, chosen by measurement on MI355X:
// FP32 planar, val[q*189 + t*63 + lane] (three dword loads, each one
// contiguous 252-byte run across the wavefront); FP64 per-lane triples,
// val[(q*63 + lane)*3 + t] (one dwordx4 + dwordx2 per lane).
template <typename T>
constexpr bool planar_vals() {
  return sizeof(T) == 4;
}

constexpr int kSliceRows = 21;
constexpr int kSliceVals = 9 * kSliceRows;  // 189
constexpr int kPackBlock = 256;              // 4 wavefronts = 4 slices per workgroup
constexpr int kSlicesPerBlock = kPackBlock / kWave;

// A group whose slices all hold the same number of block-slots K (the
// group's longest row; shorter rows padded with zero blocks of column 0) and
// whose rows are consecutive: a kernel computes a slice's first block-slot
// and its rows from these instead of loading them, which takes one memory
// round trip off the chain of dependent loads of every row.
struct GroupView {
  unsigned uni;    // 1: the group has this shape (else use the tables)
  unsigned q0, K;  // first block-slot of the group's first slice, slots per slice
  unsigned r0, nrows;
};

// Host-built index structure of a plan; values are packed per precision.
struct PackedPlan {
  unsigned nslices = 0;
  size_t nq = 0;                          // block-slots over all slices
  std::vector<unsigned> group_ptr;        // slices of group g: [group_ptr[g], group_ptr[g+1])
  DBuf<unsigned> off, kmax;               // per slice
  DBuf<unsigned> row, len;                // per slice row (row = ~0u for padding)
  DBuf<unsigned> col;                     // per block-slot x slice row
  DBuf<int> src;                          // CRS slot feeding (q, r), -1 for padding
  std::vector<GroupView> gview;           // per group
  bool built = false;

  // groups: list of row lists; ranges: CRS slot range [s0, s1) of each row's
  // blocks.  uniform: give a group of consecutive rows the GroupView shape
  // when padding every slice to the longest row adds at most 10% slots.
  void build(const std::vector<std::vector<unsigned>> &groups, const std::vector<unsigned> &s0,
             const std::vector<unsigned> &s1, const std::vector<unsigned> &crs_col,
             bool uniform = false) {
    std::vector<unsigned> h_off, h_kmax, h_row, h_len, h_col;
    std::vector<int> h_src;
    group_ptr.assign(1, 0);
    gview.clear();
    size_t q = 0;
    for (const auto &g : groups) {
      GroupView gv{0u, (unsigned)q, 0u, 0u, (unsigned)g.size()};
      if (uniform && !g.empty()) {
        bool consecutive = true;
        unsigned K = 0;
        size_t tight = 0;
        for (size_t k = 0; k < g.size(); ++k) {
          consecutive = consecutive && g[k] == g[0] + k;
          K = std::max(K, s1[g[k]] - s0[g[k]]);
        }
        for (size_t b = 0; b < g.size(); b += kSliceRows) {
          unsigned km = 0;
          for (size_t k = b; k < std::min(g.size(), b + kSliceRows); ++k)
            km = std::max(km, s1[g[k]] - s0[g[k]]);
          tight += km;
        }
        const size_t nsl = (g.size() + kSliceRows - 1) / kSliceRows;
        if (consecutive && (double)K * nsl <= 1.1 * (double)tight) {
          gv.uni = 1;
          gv.K = K;
          gv.r0 = g[0];
        }
      }
      gview.push_back(gv);
      for (size_t b = 0; b < g.size(); b += kSliceRows) {
        const size_t e = std::min(g.size(), b + kSliceRows);
        unsigned km = 0;
        for (size_t k = b; k < e; ++k) km = std::max(km, s1[g[k]] - s0[g[k]]);
        if (gv.uni) km = gv.K;
        h_off.push_back((unsigned)q);
        h_kmax.push_back(km);
        for (int rr = 0; rr < kSliceRows; ++rr) {
          const size_t k = b + rr;
          h_row.push_back(k < e ? g[k] : ~0u);
          h_len.push_back(k < e ? s1[g[k]] - s0[g[k]] : 0u);
        }
        for (unsigned kk = 0; kk < km; ++kk)
          for (int rr = 0; rr < kSliceRows; ++rr) {
            const size_t k = b + rr;
            if (k < e && kk < s1[g[k]] - s0[g[k]]) {
              const unsigned s = s0[g[k]] + kk;
              h_col.push_back(crs_col[s]);
              h_src.push_back((int)s);
            } else {
              h_col.push_back(0u);
              h_src.push_back(-1);
            }
          }
        q += km;
      }
      group_ptr.push_back((unsigned)h_off.size());
    }
    nslices = (unsigned)h_off.size();
    nq = q;
    off.upload(h_off);
    kmax.upload(h_kmax);
    row.upload(h_row);
    len.upload(h_len);
    col.upload(h_col);
    src.upload(h_src);
    built = true;
  }
  unsigned group_slices(size_t g) const { return group_ptr[g + 1] - group_ptr[g]; }
};

// Packed values of one plan in precision T.
template <typename T>
struct PackedVals {
  DBuf<T> val, dia;
  uint64_t version = ~0ull;
};

// ---------------------------------------------------------------------------
// packing kernels (gather from CRS storage, optional rounding to T)

// One thread per (block-slot or slice, lane < 63): the lane's three values
// (one row of a 3x3 block: contiguous in CRS) are read together and written
// to their places in the layout.  Launch with pack_grid(count) x 256.
inline unsigned pack_grid(size_t count) {
  return (unsigned)((count * 63 + 255) / 256);
}

template <typename T>
__device__ inline void put_lane(T *__restrict__ out, size_t q, unsigned lane, const double *v) {
  if (planar_vals<T>()) {
    T *o = out + q * kSliceVals + lane;
    o[0] = v ? (T)v[0] : (T)0;
    o[63] = v ? (T)v[1] : (T)0;
    o[126] = v ? (T)v[2] : (T)0;
  } else {
    T *o = out + (q * 63 + lane) * 3;
    o[0] = v ? (T)v[0] : (T)0;
    o[1] = v ? (T)v[1] : (T)0;
    o[2] = v ? (T)v[2] : (T)0;
  }
}

template <typename T>
__global__ void k_pack_blocks(size_t nq, const int *__restrict__ src, const double *__restrict__ crs,
                              T *__restrict__ val) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // over nq*63
  if (t >= nq * 63) return;
  const size_t q = t / 63;
  const unsigned lane = (unsigned)(t - q * 63);
  const unsigned c = lane / kSliceRows, r = lane % kSliceRows;
  const int s = src[q * kSliceRows + r];
  put_lane(val, q, lane, s >= 0 ? crs + (size_t)s * 9 + c * 3 : nullptr);
}

template <typename T>
__global__ void k_pack_dia(unsigned nslices, const unsigned *__restrict__ row,
                           const double *__restrict__ dia_crs, T *__restrict__ dia) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // over nslices*63
  if (t >= (size_t)nslices * 63) return;
  const size_t s = t / 63;
  const unsigned lane = (unsigned)(t - s * 63);
  const unsigned c = lane / kSliceRows, r = lane % kSliceRows;
  const unsigned i = row[s * kSliceRows + r];
  put_lane(dia, s, lane, i != ~0u ? dia_crs + (size_t)i * 9 + c * 3 : nullptr);
}

// Row-sum blocks of the difference form (sliced_row_diff_sum): in the
// diagonal slot of row i, S_i = D_i + sum_j A_ij over the row's CRS blocks
// [rowptr[i], rowptr[i+1]), summed in FP64 and rounded once to T.
template <typename T>
__global__ void k_pack_rowsum(unsigned nslices, const unsigned *__restrict__ row,
                              const unsigned *__restrict__ rowptr,
                              const double *__restrict__ crs, const double *__restrict__ dia_crs,
                              T *__restrict__ dia) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // over nslices*63
  if (t >= (size_t)nslices * 63) return;
  const size_t s = t / 63;
  const unsigned lane = (unsigned)(t - s * 63);
  const unsigned c = lane / kSliceRows, r = lane % kSliceRows;
  const unsigned i = row[s * kSliceRows + r];
  if (i == ~0u) {
    put_lane(dia, s, lane, (const double *)nullptr);
    return;
  }
  const double *d = dia_crs + (size_t)i * 9 + c * 3;
  double v[3] = {d[0], d[1], d[2]};
  for (unsigned q = rowptr[i]; q < rowptr[i + 1]; ++q) {
    const double *b = crs + (size_t)q * 9 + c * 3;
    v[0] += b[0];
    v[1] += b[1];
    v[2] += b[2];
  }
  put_lane(dia, s, lane, v);
}

// ---------------------------------------------------------------------------
// row kernels

struct SliceLane {
  unsigned s;     // slice
  unsigned lane;  // 0..63
  unsigned c, r;  // component, slice row (c == 3 for the idle lane)
};

// Workgroup number used by the slice kernels.  The dispatcher hands
// workgroup b to XCD b % 8, and each XCD has its own L2.  With xcd set the
// workgroups are renumbered so that every XCD runs one contiguous run of
// slices: the vector entries gathered by neighbouring rows then stay inside
// one L2 instead of being pulled into all eight, which pays while one XCD's
// share of the vector fits its L2 (measured on MI355X, 64^3 cube, 4-colour
// order: FP32 SpMV 34.5 -> 31.1 us) and costs where it does not (128^3 cube:
// the solve 5% slower), so the solver's autotuner picks it per problem.  Any
// bijection is correct; partial sums are stored at this number, so their
// order does not depend on the dispatch.
constexpr unsigned kXcds = 8;
__device__ inline unsigned wg_id(bool xcd) {
  const unsigned b = blockIdx.x;
  if (!xcd) return b;
  const unsigned g = gridDim.x, q = g / kXcds, rem = g % kXcds, x = b % kXcds;
  return x * q + min(x, rem) + b / kXcds;
}

// Workgroup number by segments: the logical workgroups form nseg segments
// [seg[k], seg[k+1]) whose sizes are multiples of 8 except the last; XCD x
// takes run x of each segment's eight equal runs (the last segment's runs
// follow the grid's own remainder, so the map is a bijection).  For a
// matrix whose rows are grouped by colour, one segment per colour gives
// every XCD the same spatial slab of every colour: the neighbours a row
// gathers then lie in its own XCD's L2.  nseg = 1 is wg_id(true).
constexpr unsigned kMaxSeg = 8;
__device__ inline unsigned seg_wg_id(const unsigned *seg, unsigned nseg) {
  const unsigned b = blockIdx.x, x = b % kXcds;
  unsigned k = b / kXcds;  // the k-th workgroup that XCD x runs
  for (unsigned s = 0; s + 1 < nseg; ++s) {
    const unsigned run = (seg[s + 1] - seg[s]) / kXcds;
    if (k < run) return seg[s] + x * run + k;
    k -= run;
  }
  const unsigned n = seg[nseg] - seg[nseg - 1], q = n / kXcds, rem = n % kXcds;
  return seg[nseg - 1] + x * q + min(x, rem) + k;
}

__device__ inline SliceLane slice_lane(unsigned slice0, bool xcd = false) {
  SliceLane L;
  L.lane = threadIdx.x & (kWave - 1);
  L.s = slice0 + wg_id(xcd) * kSlicesPerBlock + threadIdx.x / kWave;
  L.c = L.lane / kSliceRows;
  L.r = L.lane - L.c * kSliceRows;
  return L;
}

template <typename T>
struct Trip {
  T a, b, c;
};
template <typename T>
struct alignas(4 * sizeof(T)) Quad {
  T a, b, c, d;
};

// the three values of lane `lane`'s block row in block-slot q.  NT: loads
// marked non-temporal (streamed values that should not displace the
// gathered vector lines from the caches; measured: FP32 sweeps -14%, FP64
// kernels and the FP32 SpMV slower, so only the FP32 sweeps use it)
template <typename T, bool NT = false>
__device__ inline Trip<T> load_trip(const T *__restrict__ base, size_t q, unsigned lane) {
  if (planar_vals<T>()) {
    const T *p = base + q * kSliceVals + lane;
    if (NT)
      return Trip<T>{__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 63),
                     __builtin_nontemporal_load(p + 126)};
    return Trip<T>{p[0], p[63], p[126]};
  }
  const T *p = base + (q * 63 + lane) * 3;
  if (NT)
    return Trip<T>{__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1),
                   __builtin_nontemporal_load(p + 2)};
  return *reinterpret_cast<const Trip<T> *>(p);
}

// non-temporal value loads for a sweep in storage precision M
template <typename M>
constexpr bool sweep_nt() {
  return sizeof(M) == 4;
}

// node j of a vector with node stride VS
template <typename T, int VS>
__device__ inline Trip<T> load_node(const T *x, size_t j) {
  if (VS == 4) {
    const Quad<T> v = *reinterpret_cast<const Quad<T> *>(x + 4 * j);
    return Trip<T>{v.a, v.b, v.c};
  }
  return Trip<T>{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
}

// Accumulates this lane's component over the row's blocks in stored order:
//   SUB:  acc -= (v0*x0 + v1*x1) + v2*x2
//   else: acc += alpha * ((v0*x0 + v1*x1) + v2*x2)
// U blocks per round: their column indices and values are loaded first (one
// memory round trip), then the U gathers of x (a second one), then the U
// updates in order.  Values are stored in precision M and widened to the
// vector precision T before the products (M = T: no conversion).  The row
// has K blocks; slots below KL >= K are loaded (KL = K, or the slice's slot
// count when padding slots hold zero blocks of column 0: then no load waits
// for K).
template <typename T, typename M, bool SUB, int U, int VS, bool NT = false>
__device__ inline T sliced_row_sum_m(T acc, T alpha, const SliceLane &L, unsigned K, unsigned KL,
                                     unsigned kmax, unsigned q0, const unsigned *__restrict__ col,
                                     const M *__restrict__ val, const T *x) {
  for (unsigned k0 = 0; k0 < kmax; k0 += U) {
    unsigned j[U];
    Trip<M> v[U];
    Trip<T> xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u < KL) {
        const size_t q = q0 + k0 + u;
        j[u] = col[q * kSliceRows + L.r];
        v[u] = load_trip<M, NT>(val, q, L.lane);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u < KL) xv[u] = load_node<T, VS>(x, j[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u < K) {
        const T d = (T)v[u].a * xv[u].a + (T)v[u].b * xv[u].b + (T)v[u].c * xv[u].c;
        if (SUB)
          acc -= d;
        else
          acc += alpha * d;
      }
    }
  }
  return acc;
}

// Difference form of a row sum (the mixed solver's FP32 copy of A):
//   acc + sum_k ((v0*(x0-o0) + v1*(x1-o1)) + v2*(x2-o2))
// over the row's off-diagonal blocks, o = the row's own node values.  The
// diagonal slot then holds the row-sum block S_i = D_i + sum_j A_ij
// (k_pack_rowsum), and acc + S_i o is (A x)_i.  A stiffness matrix maps
// rigid translations to zero, so on smooth vectors (A x)_i is a small
// difference of large products: in this form the rounding of the stored
// blocks to FP32 multiplies x_j - x_i instead of x_j, and S_i (near zero
// away from fixed nodes) is rounded once from its FP64 sum.  Same loads and
// load rounds as sliced_row_sum_m.
template <typename T, typename M, int U, int VS>
__device__ inline T sliced_row_diff_sum(T acc, const Trip<T> &o, const SliceLane &L, unsigned K,
                                        unsigned KL, unsigned kmax, unsigned q0,
                                        const unsigned *__restrict__ col,
                                        const M *__restrict__ val, const T *x) {
  for (unsigned k0 = 0; k0 < kmax; k0 += U) {
    unsigned j[U];
    Trip<M> v[U];
    Trip<T> xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u < KL) {
        const size_t q = q0 + k0 + u;
        j[u] = col[q * kSliceRows + L.r];
        v[u] = load_trip<M>(val, q, L.lane);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u < KL) xv[u] = load_node<T, VS>(x, j[u]);
    // the differences in a pass of their own, after every gather of the
    // round is requested (formed inside each product, the compiler issued
    // the gathers one at a time)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u < K) {
        xv[u].a -= o.a;
        xv[u].b -= o.b;
        xv[u].c -= o.c;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u < K)
        acc += (T)v[u].a * xv[u].a + (T)v[u].b * xv[u].b + (T)v[u].c * xv[u].c;
  }
  return acc;
}

// One load round of a row sum: column indices and value triples of the
// slots k0, k0 + H, ..., k0 + (U - 1) H that lie below KL.
template <typename M, int U>
struct SlotRound {
  unsigned j[U];
  Trip<M> v[U];
};

template <typename M, int U, int H, bool NT = false>
__device__ inline void load_slot_round(SlotRound<M, U> &R, const SliceLane &L, unsigned k0,
                                       unsigned KL, unsigned q0,
                                       const unsigned *__restrict__ col,
                                       const M *__restrict__ val) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const unsigned k = k0 + H * u;
    if (k < KL) {
      const size_t q = q0 + k;
      R.j[u] = col[q * kSliceRows + L.r];
      R.v[u] = load_trip<M, NT>(val, q, L.lane);
    }
  }
}

// Gathers and products of one loaded round (slots k0 + H u), accumulated
// into acc as sliced_row_sum_m does.
template <typename T, typename M, bool SUB, int U, int H, int VS>
__device__ inline T slot_round_sum(T acc, T alpha, const SlotRound<M, U> &R, unsigned k0,
                                   unsigned K, unsigned KL, const T *x) {
  Trip<T> xv[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (k0 + H * u < KL) xv[u] = load_node<T, VS>(x, R.j[u]);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (k0 + H * u < K) {
      const T d = (T)R.v[u].a * xv[u].a + (T)R.v[u].b * xv[u].b + (T)R.v[u].c * xv[u].c;
      if (SUB)
        acc -= d;
      else
        acc += alpha * d;
    }
  }
  return acc;
}

// sliced_row_sum_m over the slots h, h + H, h + 2H, ... (part h of a row
// split into H interleaved parts; H = 1: the whole row), whose first round
// the caller has already requested with load_slot_round(R0, L, h, ...) --
// so it can issue those loads before anything they need not wait for.  Same
// products, same order as sliced_row_sum_m.  Later rounds load into fresh
// registers (a round carried through the loop makes the compiler wait for
// every outstanding load before each reload).
template <typename T, typename M, bool SUB, int U, int H, int VS, bool NT = false>
__device__ inline T sliced_row_sum_pre(T acc, T alpha, const SlotRound<M, U> &R0,
                                       const SliceLane &L, unsigned h, unsigned K, unsigned KL,
                                       unsigned kmax, unsigned q0,
                                       const unsigned *__restrict__ col,
                                       const M *__restrict__ val, const T *x) {
  if (h < kmax) acc = slot_round_sum<T, M, SUB, U, H, VS>(acc, alpha, R0, h, K, KL, x);
  for (unsigned k0 = h + H * U; k0 < kmax; k0 += H * U) {
    SlotRound<M, U> R;
    load_slot_round<M, U, H, NT>(R, L, k0, KL, q0, col, val);
    acc = slot_round_sum<T, M, SUB, U, H, VS>(acc, alpha, R, k0, K, KL, x);
  }
  return acc;
}

template <typename T, bool SUB, int U, int VS>
__device__ inline T sliced_row_sum(T acc, T alpha, const SliceLane &L, unsigned K, unsigned kmax,
                                   unsigned q0, const unsigned *__restrict__ col,
                                   const T *__restrict__ val, const T *x) {
  return sliced_row_sum_m<T, T, SUB, U, VS>(acc, alpha, L, K, K, kmax, q0, col, val, x);
}

// this lane's row of a stored 3x3 block (precision M) times (x0, x1, x2)
template <typename T, typename M, bool NT = false>
__device__ inline T dia_row_dot_m(const M *__restrict__ dia, const SliceLane &L, T x0, T x1,
                                  T x2) {
  const Trip<M> d = load_trip<M, NT>(dia, L.s, L.lane);
  return (T)d.a * x0 + (T)d.b * x1 + (T)d.c * x2;
}

template <typename T>
__device__ inline T dia_row_dot(const T *__restrict__ dia, const SliceLane &L, T x0, T x1, T x2) {
  return dia_row_dot_m<T, T>(dia, L, x0, x1, x2);
}

struct PlanView {
  const unsigned *off, *kmax, *row, *len, *col;
};

inline PlanView view(const PackedPlan &P) {
  return PlanView{P.off.p, P.kmax.p, P.row.p, P.len.p, P.col.p};
}

// Lane L's row in a slice of a group whose first slice is slice0 (sl: the
// slice exists), and the slice's slot range: computed for a GroupView-shaped
// group, loaded from the tables otherwise.  i = ~0u: no row (idle lane,
// padding row).  kl: slots the row sum may load without waiting for len.
struct RowSlots {
  unsigned i, len, kl, kmax, off;
};

__device__ inline unsigned row_of(const PlanView &P, const GroupView &g, const SliceLane &L,
                                  unsigned slice0, bool sl) {
  if (!sl || L.c >= 3) return ~0u;
  if (!g.uni) return P.row[L.s * kSliceRows + L.r];
  const unsigned rr = (L.s - slice0) * kSliceRows + L.r;
  return rr < g.nrows ? g.r0 + rr : ~0u;
}

__device__ inline RowSlots row_slots(const PlanView &P, const GroupView &g, const SliceLane &L,
                                     unsigned slice0, bool sl) {
  RowSlots R{~0u, 0u, 0u, 0u, 0u};
  if (g.uni) {
    const unsigned ls = L.s - slice0, rr = ls * kSliceRows + L.r;
    if (sl && L.c < 3 && rr < g.nrows) {
      R.i = g.r0 + rr;
      R.len = P.len[L.s * kSliceRows + L.r];
      R.kl = g.K;
      R.kmax = g.K;
      R.off = g.q0 + ls * g.K;
    }
  } else if (sl && L.c < 3) {
    R.i = P.row[L.s * kSliceRows + L.r];
    R.len = P.len[L.s * kSliceRows + L.r];  // 0 for a padding row
    R.kl = R.len;
    R.kmax = P.kmax[L.s];
    R.off = P.off[L.s];
  }
  return R;
}

// blocks per load round of the plain MatVec (measured on MI355X, 64^3
// cube, 4-colour order: 4 for FP64 / 8 for FP32)
template <typename T>
constexpr int unroll_spmv() {
  return sizeof(T) == 8 ? 4 : 8;
}

// y = alpha*A*x + beta*y over the slices [slice0, slice0+nsl) (all rows)
template <typename T, int VS>
__global__ void __launch_bounds__(kPackBlock)
k_sliced_matvec(unsigned slice0, unsigned nsl, PlanView P, const T *__restrict__ val,
                const T *__restrict__ dia, T *__restrict__ y, T alpha, const T *__restrict__ x,
                T beta) {
  const SliceLane L = slice_lane(slice0);
  if (L.s >= slice0 + nsl || L.c >= 3) return;
  const unsigned i = P.row[L.s * kSliceRows + L.r];
  if (i == ~0u) return;
  T acc = y[VS * (size_t)i + L.c] * beta;
  acc = sliced_row_sum<T, false, unroll_spmv<T>(), VS>(
      acc, alpha, L, P.len[L.s * kSliceRows + L.r], P.kmax[L.s], P.off[L.s], P.col, val, x);
  const Trip<T> xi = load_node<T, VS>(x, i);
  acc += alpha * dia_row_dot(dia, L, xi.a, xi.b, xi.c);
  y[VS * (size_t)i + L.c] = acc;
}

// forward sweep of one level: vec_i = D_i^-1 (src_i - sum L_ij vec_j), over
// the level's slices [slice0, slice0 + nsl) with its group shape gv (the
// row and slot offsets computed instead of loaded when the group is
// uniform); the row's own entry, its D^-1 row and the first round of
// blocks are requested together.  done (nullable): the kernel returns at
// once when *done is set (a fused solver's convergence flag; read first, so
// that launches after convergence cost only their launch).  Products and
// order of sliced_row_sum_m.  4 blocks per load round (measured at 64^3:
// SolvePrecond 95.4 us with the previous form, 93.1 with 8, 91.4 with 4).
template <typename T, int VS>
__global__ void __launch_bounds__(kPackBlock)
k_sliced_fwd(unsigned slice0, unsigned nsl, PlanView P, GroupView gv, const T *__restrict__ val,
             const T *__restrict__ dia, const T *src, T *vec, const int *done) {
  constexpr int U = 4;
  if (done && ld_state(done)) return;
  const SliceLane L = slice_lane(slice0);
  const bool sl = L.s < slice0 + nsl;
  if (!sl) return;  // whole wavefront
  const RowSlots R = row_slots(P, gv, L, slice0, sl);
  const unsigned i = R.i;
  T t = 0;
  Trip<T> d{};
  SlotRound<T, U> R0;
  if (i != ~0u) {
    t = src[VS * (size_t)i + L.c];
    d = load_trip<T>(dia, L.s, L.lane);
    load_slot_round<T, U, 1>(R0, L, 0, R.kl, R.off, P.col, val);
  }
  if (i != ~0u)
    t = sliced_row_sum_pre<T, T, true, U, 1, VS>(t, (T)1, R0, L, 0, R.len, R.kl, R.kmax, R.off,
                                                 P.col, val, vec);
  const T t0 = __shfl(t, (int)L.r, kWave);
  const T t1 = __shfl(t, (int)L.r + kSliceRows, kWave);
  const T t2 = __shfl(t, (int)L.r + 2 * kSliceRows, kWave);
  if (i != ~0u) vec[VS * (size_t)i + L.c] = d.a * t0 + d.b * t1 + d.c * t2;
}

// backward sweep of one level: vec_i -= sum U_ij vec_j (U scaled by D^-1);
// same structure as k_sliced_fwd
template <typename T, int VS>
__global__ void __launch_bounds__(kPackBlock)
k_sliced_bwd(unsigned slice0, unsigned nsl, PlanView P, GroupView gv, const T *__restrict__ val,
             T *vec, const int *done) {
  constexpr int U = 4;
  if (done && ld_state(done)) return;
  const SliceLane L = slice_lane(slice0);
  const bool sl = L.s < slice0 + nsl;
  if (!sl) return;  // whole wavefront
  const RowSlots R = row_slots(P, gv, L, slice0, sl);
  const unsigned i = R.i;
  if (i == ~0u) return;
  T t = vec[VS * (size_t)i + L.c];
  SlotRound<T, U> R0;
  load_slot_round<T, U, 1>(R0, L, 0, R.kl, R.off, P.col, val);
  t = sliced_row_sum_pre<T, T, true, U, 1, VS>(t, (T)1, R0, L, 0, R.len, R.kl, R.kmax, R.off,
                                               P.col, val, vec);
  vec[VS * (size_t)i + L.c] = t;
}

inline unsigned slice_grid(unsigned nsl) { return std::max(1u, ceil_div(nsl, kSlicesPerBlock)); }

}  // namespace dfm2h
