// handles.hpp -- the opaque objects behind the C ABI handles.
//
// Values are kept twice in HBM: in the reference's CRS layout (what
// assembly, fixed BCs, CopyValue, Decompose and upload/download work on) and
// packed into wavefront slices (what every iteration kernel streams).  A
// packed copy remembers the value version it was made from and is rebuilt
// lazily when the CRS values changed.
#pragma once

#include "packed.hpp"

namespace dfm2h {

struct Pattern {
  unsigned nblk = 0;
  size_t ncrs = 0;
  std::vector<unsigned> colind, rowptr;  // host copies
  DBuf<unsigned> colind_d, rowptr_d;
  DBuf<unsigned> rowof_d;  // block row of each crs slot
};

struct PcgWork;  // pcg.hip
void destroy_pcg_work(PcgWork *);

// Applies the matrix's staged per-element Merge additions (bsr.hip); every
// entry point that reads or writes the matrix's values calls it first.
void flush_merge(dfm2h_bsr_s *A);
// whether A is a matrix handle that has been created and not destroyed
bool bsr_alive(const dfm2h_bsr_s *A);

struct BicgWork;  // krylov.hip
void destroy_bicg_work(BicgWork *);

// ILU(0) sweeps on FP64 vectors, level by level on stream s (ilu.hip):
// forward vec = L^-1 src (src may be vec), backward vec = U^-1 vec; with
// done set the kernels return at once when *done is nonzero.
void ilu_forward_levels_double(dfm2h_ilu_s *P, const double *src, double *vec, hipStream_t s,
                               const int *done = nullptr);
void ilu_backward_levels_double(dfm2h_ilu_s *P, double *vec, hipStream_t s,
                                const int *done = nullptr);

// Packs CRS values (and optional per-row diagonal blocks) into V when the
// source version differs from the packed one.
template <typename T>
void ensure_packed(const PackedPlan &P, PackedVals<T> &V, const double *crs, const double *dia,
                   uint64_t version, hipStream_t s) {
  if (V.version == version && (V.val.p || P.nq == 0)) return;
  V.val.alloc(P.nq * kSliceVals);
  if (P.nq) {
    hipLaunchKernelGGL(k_pack_blocks<T>, dim3(pack_grid(P.nq)), dim3(256), 0, s, P.nq, P.src.p,
                       crs, V.val.p);
  }
  if (dia) {
    V.dia.alloc((size_t)P.nslices * kSliceVals);
    if (P.nslices) {
      hipLaunchKernelGGL(k_pack_dia<T>, dim3(pack_grid(P.nslices)), dim3(256), 0, s, P.nslices,
                         P.row.p, dia, V.dia.p);
    }
  }
  check_launch("k_pack");
  V.version = version;
}

// The same copy in the difference form of the row sums: off-diagonal blocks
// as ensure_packed, the row-sum blocks D_i + sum_j A_ij (rowptr: the CRS row
// starts, on the device) in the diagonal slot.
template <typename T>
void ensure_packed_rowsum(const PackedPlan &P, PackedVals<T> &V, const double *crs,
                          const double *dia, const unsigned *rowptr, uint64_t version,
                          hipStream_t s) {
  if (V.version == version && (V.val.p || P.nq == 0)) return;
  V.val.alloc(P.nq * kSliceVals);
  if (P.nq) {
    hipLaunchKernelGGL(k_pack_blocks<T>, dim3(pack_grid(P.nq)), dim3(256), 0, s, P.nq, P.src.p,
                       crs, V.val.p);
  }
  V.dia.alloc((size_t)P.nslices * kSliceVals);
  if (P.nslices) {
    hipLaunchKernelGGL(k_pack_rowsum<T>, dim3(pack_grid(P.nslices)), dim3(256), 0, s, P.nslices,
                       P.row.p, rowptr, crs, dia, V.dia.p);
  }
  check_launch("k_pack_rowsum");
  V.version = version;
}

// Groups rows by dependency level.  Forward: lev(i) = 1 + max lev(j) over
// the lower part of row i (0 when empty); backward: the same over the upper
// part, rows visited from the last one.  Rows inside a level stay ascending.
inline void build_levels(unsigned n, const std::vector<unsigned> &colind,
                         const std::vector<unsigned> &rowptr, const std::vector<unsigned> &diaind,
                         bool forward, std::vector<unsigned> &levptr,
                         std::vector<unsigned> &levrows) {
  std::vector<unsigned> lev(n, 0);
  unsigned nlev = 0;
  if (forward) {
    for (unsigned i = 0; i < n; ++i) {
      unsigned l = 0;
      for (unsigned s = colind[i]; s < diaind[i]; ++s) l = std::max(l, lev[rowptr[s]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (unsigned i = n; i-- > 0;) {
      unsigned l = 0;
      for (unsigned s = diaind[i]; s < colind[i + 1]; ++s) l = std::max(l, lev[rowptr[s]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  levptr.assign(nlev + 1, 0);
  for (unsigned i = 0; i < n; ++i) levptr[lev[i] + 1]++;
  for (unsigned l = 0; l < nlev; ++l) levptr[l + 1] += levptr[l];
  levrows.resize(n);
  std::vector<unsigned> cur(levptr.begin(), levptr.end() - 1);
  for (unsigned i = 0; i < n; ++i) levrows[cur[lev[i]]++] = i;
}

}  // namespace dfm2h

struct dfm2h_bsr_s {
  unsigned nblk = 0;
  bool has_pattern = false;
  dfm2h::Pattern pat;
  dfm2h::DBuf<double> vcrs, vdia;
  uint64_t version = 1;  // bumped whenever values change
  uint64_t pattern_serial = 0;  // process-unique, set with the pattern (next_serial)
  // Per-element Merge (dfm2h_bsr_merge): additions staged on the host in
  // call order -- destination (slot < ncrs: off-diagonal slot; ncrs + i:
  // diagonal block i) and the 3x3 block -- and applied on the device, in
  // that order per destination, before the values are next used
  std::vector<unsigned> mg_dst;
  std::vector<double> mg_val;   // 9 per staged block
  std::vector<unsigned> mg_mark;  // per column block: its slot in the current row, or ~0u
  dfm2h::DBuf<double> mg_val_d;
  dfm2h::DBuf<unsigned> mg_run_d;  // per touched destination: {dst, first, count}
  dfm2h::PackedPlan plan;  // all rows, CRS block order
  dfm2h::PackedVals<double> pk64;
  dfm2h::PackedVals<float> pk32;
  // gather assembly: per destination block (off-diagonal slots, then the
  // diagonal blocks) the entries (element | li << 28 | lj << 30) of the
  // elements it receives, ascending; built on first use with a mesh
  uint64_t asm_mesh_serial = 0;  // serial of that mesh (0: none yet)
  int asm_missing = 0;           // an element edge lacks a slot in the pattern
  dfm2h::DBuf<unsigned> asm_ptr, asm_ent;
  dfm2h::DBuf<double> asm_rec;  // per-tet records (geometry, Lame), rewritten every call
};

struct dfm2h_tetmesh_s {
  uint64_t serial = 0;  // unique per created mesh
  size_t ntet = 0, np = 0;
  unsigned max_elsup = 0;           // most elements around one node
  // nodes with more elements than the LDS form of the device pattern build
  // takes, and where each one's candidate run starts (4 per element)
  size_t nbig = 0, big_cand = 0;
  dfm2h::DBuf<unsigned> big, big_off;
  dfm2h::DBuf<unsigned> tet;        // [4*ntet]
  dfm2h::DBuf<unsigned> elsup_ind;  // [np+1]
  dfm2h::DBuf<unsigned> elsup;      // element ids, ascending per node
};

struct dfm2h_ilu_s {
  unsigned nblk = 0;
  size_t ncrs = 0;
  bool has_pattern = false;
  bool decomposed = false;
  std::vector<unsigned> colind, rowptr, diaind;  // host (sorted pattern)
  dfm2h::DBuf<unsigned> colind_d, rowptr_d, diaind_d;
  dfm2h::DBuf<double> vcrs, vdia;
  // A slot -> ILU slot map for CopyValue, built against the matrix
  // `mapped_for` with pattern serial `mapped_serial`
  const dfm2h_bsr_s *mapped_for = nullptr;
  uint64_t mapped_serial = 0;
  dfm2h::DBuf<int> a2l_d;
  bool a2l_identity = false;  // a2l[s] == s for every slot (same pattern)
  // level schedules: rows grouped by dependency level, ascending inside a level
  std::vector<unsigned> levptr_f, levptr_b;
  dfm2h::DBuf<unsigned> levrows_f;
  bool flev_ranges = false;  // every forward level is a consecutive, ascending row range
  dfm2h::DBuf<int> nsing_d;
  dfm2h::DBuf<unsigned char> sing_d;  // per row: the pivot was singular
  // the matrix the last CopyValue read, its pattern serial and value version
  const dfm2h_bsr_s *cv_from = nullptr;
  uint64_t cv_serial = 0, cv_version = 0;
  unsigned dec_lanes = 0;  // lanes per row of the factorisation kernel (0: one thread per row)
  // sliced sweeps: forward plan (lower blocks + D^-1), backward plan (upper blocks)
  dfm2h::PackedPlan fplan, bplan;
  // per forward slice: bit r set when slice row r has no upper blocks (its
  // forward result is final); per forward level: whether it has such rows
  dfm2h::DBuf<unsigned> fumask;
  std::vector<char> flev_u0;
  dfm2h::PackedVals<double> f64, b64;
  dfm2h::PackedVals<float> f32, b32;
  uint64_t version = 1;
  uint64_t pattern_serial = 0;  // process-unique, set with the pattern (next_serial)
  dfm2h::PcgWork *work = nullptr;
  dfm2h::BicgWork *bicg = nullptr;  // the BiCGStab solver's buffers and graphs
  ~dfm2h_ilu_s() {
    dfm2h::destroy_pcg_work(work);
    dfm2h::destroy_bicg_work(bicg);
  }
};
