// runtime.hip -- device selection, stream, error strings, memory helpers,
// host pattern build and the tet connectivity handle.
#include "handles.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace dfm2h {

static thread_local std::string g_last_error;
// The stream of the calling thread: a thread that sets one does not
// redirect another thread's calls, so separate handles can be driven from
// separate threads on separate streams.
static thread_local hipStream_t g_stream = nullptr;

void set_error(const std::string &msg) { g_last_error = msg; }
hipStream_t stream() { return g_stream; }

uint64_t next_serial() {
  static std::atomic<uint64_t> serial{0};
  return ++serial;
}

void require_device(const void *p, const char *what) {
  if (!p) return;
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear the failure so later launch checks do not see it
    throw Error(DFM2H_ERR_INVALID, std::string(what) +
                                       " is not device memory (a host pointer where a "
                                       "device pointer is required)");
  }
  const bool ok = at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged ||
                  at.type == hipMemoryTypeUnified ||
                  (at.type == hipMemoryTypeHost && at.devicePointer != nullptr);
  if (!ok)
    throw Error(DFM2H_ERR_INVALID, std::string(what) +
                                       " is not device memory (a host pointer where a "
                                       "device pointer is required)");
}

void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(DFM2H_ERR_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(e));
}

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<std::string, int, int>, int> g_tune;
bool g_tune_loaded = false;

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char *f = std::getenv("DFM2H_TUNE_CACHE");
  if (!f) return;
  FILE *fp = std::fopen(f, "r");
  if (!fp) return;
  char kind[32];
  int b, p, v;
  double ms;
  while (std::fscanf(fp, "%31s %d %d %d %lf", kind, &b, &p, &v, &ms) == 5)
    g_tune[std::make_tuple(std::string(kind), b, p)] = v;
  std::fclose(fp);
}
}  // namespace

int tune_lookup(const char *kind, int bucket, int fp32, int n_variants) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  auto it = g_tune.find(std::make_tuple(std::string(kind), bucket, fp32));
  if (it == g_tune.end() || it->second < 0 || it->second >= n_variants) return -1;
  return it->second;
}

void tune_store(const char *kind, int bucket, int fp32, int variant, double ms) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load_locked();
  g_tune[std::make_tuple(std::string(kind), bucket, fp32)] = variant;
  if (const char *f = std::getenv("DFM2H_TUNE_CACHE")) {
    if (FILE *fp = std::fopen(f, "a")) {
      std::fprintf(fp, "%s %d %d %d %.6f\n", kind, bucket, fp32, variant, ms);
      std::fclose(fp);
    }
  }
}

int dof_bucket(size_t ndof) { return ndof < 1 ? 0 : (int)std::floor(std::log2((double)ndof)); }

// ---------------------------------------------------------------------------
// one-ring pattern on the device: one thread per node gathers the nodes of
// its surrounding elements (other than itself; a degenerate element can
// repeat a node) into a candidate buffer, then emits the distinct ones in
// ascending order by repeated selection -- the same rows as
// JArray_PSuP_MeshElem + JArray_Sort.  Nodes with at most kPsupMaxElem
// elements keep their candidates in LDS; the others (any count) in a global
// buffer, one run of 4 x their element count per such node.  Pass 1
// (psup_ind == nullptr) writes each row's count into out; pass 2 writes the
// row at out[psup_ind[i]].

constexpr int kPsupMaxElem = 32;  // Kuhn-split voxel meshes: 24
constexpr int kPsupCand = 3 * kPsupMaxElem;
constexpr int kPsupBlock = 64;

__device__ inline void psup_row(unsigned i, unsigned *c, const unsigned *__restrict__ tet,
                                const unsigned *__restrict__ esi,
                                const unsigned *__restrict__ es,
                                const unsigned *__restrict__ psup_ind, unsigned *__restrict__ out) {
  unsigned nc = 0;
  for (unsigned k = esi[i]; k < esi[i + 1]; ++k) {
    const unsigned e = es[k];
    for (int a = 0; a < 4; ++a) {
      const unsigned v = tet[4 * (size_t)e + a];
      if (v != i) c[nc++] = v;
    }
  }
  unsigned cnt = 0, last = 0;
  for (;;) {
    unsigned best = 0xffffffffu;
    for (unsigned q = 0; q < nc; ++q) {
      const unsigned v = c[q];
      if ((cnt == 0 || v > last) && v < best) best = v;
    }
    if (best == 0xffffffffu) break;
    if (psup_ind) out[psup_ind[i] + cnt] = best;
    ++cnt;
    last = best;
  }
  if (!psup_ind) out[i] = cnt;
}

__global__ void __launch_bounds__(kPsupBlock)
k_psup_rows(unsigned np, const unsigned *__restrict__ tet, const unsigned *__restrict__ esi,
            const unsigned *__restrict__ es, const unsigned *__restrict__ psup_ind,
            unsigned *__restrict__ out) {
  __shared__ unsigned cand[kPsupBlock][kPsupCand + 1];
  const unsigned i = blockIdx.x * kPsupBlock + threadIdx.x;
  if (i >= np || esi[i + 1] - esi[i] > (unsigned)kPsupMaxElem) return;
  psup_row(i, cand[threadIdx.x], tet, esi, es, psup_ind, out);
}

// the nodes big[0 .. nbig) with more elements; node big[b]'s candidates at
// cand + coff[b] (4 per element around it: enough for the 3 others)
__global__ void __launch_bounds__(kPsupBlock)
k_psup_rows_big(unsigned nbig, const unsigned *__restrict__ big,
                const unsigned *__restrict__ coff, unsigned *__restrict__ cand,
                const unsigned *__restrict__ tet, const unsigned *__restrict__ esi,
                const unsigned *__restrict__ es, const unsigned *__restrict__ psup_ind,
                unsigned *__restrict__ out) {
  const unsigned b = blockIdx.x * kPsupBlock + threadIdx.x;
  if (b >= nbig) return;
  psup_row(big[b], cand + coff[b], tet, esi, es, psup_ind, out);
}

}  // namespace dfm2h

using namespace dfm2h;

// plain 16-byte copy, one vector per thread: the measured copy rate the
// kernels' bandwidth is compared with (a probe of grid sizes and of several
// vectors per thread found this form the fastest on MI355X)
__global__ void k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

extern "C" {

int dfm2h_init(int device) {
  return guard(__func__, [&] { DFM2H_HIP(hipSetDevice(device)); });
}

int dfm2h_set_stream(void *s) {
  g_stream = static_cast<hipStream_t>(s);
  return DFM2H_OK;
}

void *dfm2h_get_stream(void) { return g_stream; }

int dfm2h_synchronize(void) {
  return guard(__func__, [&] { DFM2H_HIP(hipStreamSynchronize(stream())); });
}

const char *dfm2h_last_error(void) { return g_last_error.c_str(); }

int dfm2h_measure_copy(size_t bytes, unsigned reps, double *gbs) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(gbs && bytes >= 16 && reps >= 1, DFM2H_ERR_INVALID, "bad argument");
    const size_t n = bytes / 16;
    DBuf<float4> a, b;
    a.alloc(n);
    b.alloc(n);
    DFM2H_HIP(hipMemsetAsync(a.p, 0, n * 16, stream()));
    TimingEvents<2> ev;
    const dim3 grid(ceil_div(n, 256));
    hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, stream(), a.p, b.p, n);  // warm-up
    DFM2H_HIP(hipEventRecord(ev[0], stream()));
    for (unsigned r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, stream(), a.p, b.p, n);
    DFM2H_HIP(hipEventRecord(ev[1], stream()));
    check_launch("k_copy16");
    DFM2H_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    DFM2H_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    DFM2H_REQUIRE(ms > 0.f, DFM2H_ERR_INVALID, "copy too short to time");
    *gbs = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  });
}

int dfm2h_malloc(void **p, size_t bytes) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(p, DFM2H_ERR_INVALID, "null output");
    DFM2H_HIP(hipMalloc(p, bytes));
  });
}

int dfm2h_free(void *p) {
  return guard(__func__, [&] { DFM2H_HIP(hipFree(p)); });
}

int dfm2h_memcpy_htod(void *dst, const void *src, size_t bytes) {
  return guard(__func__, [&] {
    DFM2H_DEVICE_ARG(dst);
    DFM2H_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  return guard(__func__, [&] {
    DFM2H_DEVICE_ARG(src);
    DFM2H_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  return guard(__func__, [&] {
    DFM2H_DEVICE_ARG(dst);
    DFM2H_DEVICE_ARG(src);
    DFM2H_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  });
}

// Element-around-point lists by counting sort (ascending element order per
// point), then the distinct neighbours of each point in visiting order, then
// an ascending sort of every row -- the set the reference builds with
// JArray_ElSuP_MeshElem / JArrayPointSurPoint_MeshOneRingNeighborhood /
// JArray_Sort (msh_topology_uniform.cpp:19-51, 250-316; jagarray.cpp:72-92).
int dfm2h_psup_tet(const unsigned *tet, size_t ntet, size_t np, unsigned *psup_ind,
                   unsigned *psup, size_t psup_cap, size_t *npsup) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(tet || ntet == 0, DFM2H_ERR_INVALID, "null tet");
    DFM2H_REQUIRE(psup_ind, DFM2H_ERR_INVALID, "null psup_ind");
    for (size_t i = 0; i < ntet * 4; ++i)
      DFM2H_REQUIRE(tet[i] < np, DFM2H_ERR_INVALID, "tet node index out of range");
    std::vector<unsigned> es_ind(np + 1, 0), es(ntet * 4);
    for (size_t i = 0; i < ntet * 4; ++i) es_ind[tet[i] + 1]++;
    for (size_t p = 0; p < np; ++p) es_ind[p + 1] += es_ind[p];
    {
      std::vector<unsigned> cur(es_ind.begin(), es_ind.end() - 1);
      for (size_t e = 0; e < ntet; ++e)
        for (int k = 0; k < 4; ++k) es[cur[tet[e * 4 + k]]++] = (unsigned)e;
    }
    std::vector<unsigned> flg(np, UINT_MAX);
    std::vector<unsigned> out;
    out.reserve(np * 15);
    psup_ind[0] = 0;
    for (size_t p = 0; p < np; ++p) {
      flg[p] = (unsigned)p;
      const size_t row0 = out.size();
      for (unsigned s = es_ind[p]; s < es_ind[p + 1]; ++s) {
        const unsigned e = es[s];
        for (int k = 0; k < 4; ++k) {
          const unsigned q = tet[e * 4 + k];
          if (flg[q] != p) {
            flg[q] = (unsigned)p;
            out.push_back(q);
          }
        }
      }
      std::sort(out.begin() + row0, out.end());
      psup_ind[p + 1] = (unsigned)out.size();
    }
    if (npsup) *npsup = out.size();
    if (psup) {
      DFM2H_REQUIRE(psup_cap >= out.size(), DFM2H_ERR_INVALID, "psup capacity too small");
      std::memcpy(psup, out.data(), sizeof(unsigned) * out.size());
    }
  });
}

// Multicolour node order for the level schedule: every node gets a colour
// that none of its neighbours has, and new2old lists the nodes colour by
// colour (ascending old index inside a colour), so that in the relabelled
// mesh no row has a lower neighbour of its own colour and the ILU(0) sweeps
// have at most ncolor dependency levels.  Colours by first fit in DSATUR
// order (Brelaz 1979) -- 4 on a Kuhn-split voxel grid at every size tried
// (16^3 to 64^3), the least a tet mesh allows (a tet is a 4-clique).
// O(np + npsup) with bucketed saturation counts.
int dfm2h_multicolor_order(size_t np, const unsigned *psup_ind, const unsigned *psup,
                           unsigned *new2old, unsigned *ncolor) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(psup_ind && ncolor && (new2old || np == 0), DFM2H_ERR_INVALID,
                  "null argument");
    DFM2H_REQUIRE(np < UINT_MAX, DFM2H_ERR_INVALID, "too many nodes");
    DFM2H_REQUIRE(psup_ind[0] == 0, DFM2H_ERR_INVALID, "psup_ind[0] must be 0");
    for (size_t p = 0; p < np; ++p)
      DFM2H_REQUIRE(psup_ind[p + 1] >= psup_ind[p], DFM2H_ERR_INVALID,
                    "psup_ind must be non-decreasing");
    DFM2H_REQUIRE(psup || psup_ind[np] == 0, DFM2H_ERR_INVALID, "null psup");
    // the pattern must be symmetric, without self or repeated entries (as
    // tet meshes give it): a colouring of the listed pairs is then one of the
    // rows.  Checked against its transpose (the nodes that list each node).
    const size_t m = psup_ind[np];
    std::vector<unsigned> tind(np + 1, 0), tadj(m), stamp(np, UINT_MAX);
    for (size_t k = 0; k < m; ++k) {
      DFM2H_REQUIRE(psup[k] < np, DFM2H_ERR_INVALID, "psup entry out of range");
      tind[psup[k] + 1]++;
    }
    for (size_t p = 0; p < np; ++p) tind[p + 1] += tind[p];
    {
      std::vector<unsigned> cur(tind.begin(), tind.end() - 1);
      for (size_t p = 0; p < np; ++p)
        for (unsigned k = psup_ind[p]; k < psup_ind[p + 1]; ++k) tadj[cur[psup[k]]++] = (unsigned)p;
    }
    for (size_t p = 0; p < np; ++p) {
      DFM2H_REQUIRE(tind[p + 1] - tind[p] == psup_ind[p + 1] - psup_ind[p], DFM2H_ERR_INVALID,
                    "psup is not symmetric");
      for (unsigned k = psup_ind[p]; k < psup_ind[p + 1]; ++k) {
        const unsigned q = psup[k];
        DFM2H_REQUIRE(q != p && stamp[q] != p, DFM2H_ERR_INVALID,
                      "psup has a self or repeated entry");
        stamp[q] = (unsigned)p;
      }
      for (unsigned k = tind[p]; k < tind[p + 1]; ++k)
        DFM2H_REQUIRE(stamp[tadj[k]] == p, DFM2H_ERR_INVALID, "psup is not symmetric");
    }
    // DSATUR order: next the uncoloured node whose neighbours show the most
    // distinct colours (buckets by that count, stale entries skipped: a node
    // is pushed again each time its count grows; ties go to the node pushed
    // last, at the start to the lowest index); each node takes the lowest
    // colour none of its neighbours has
    unsigned dmax = 0;
    for (size_t p = 0; p < np; ++p) dmax = std::max(dmax, psup_ind[p + 1] - psup_ind[p]);
    const size_t words = (dmax + 1 + 63) / 64;  // colours stay below dmax + 1
    std::vector<uint64_t> used(np * words, 0);
    std::vector<unsigned> nsat(np, 0), color(np, UINT_MAX);
    std::vector<std::vector<unsigned>> bucket(dmax + 2);
    bucket[0].reserve(np);
    for (size_t p = np; p-- > 0;) bucket[0].push_back((unsigned)p);
    size_t top = 0, done = 0;
    unsigned nc = 0;
    while (done < np) {
      while (bucket[top].empty()) {
        DFM2H_REQUIRE(top > 0, DFM2H_ERR_INVALID, "colouring lost a node");
        --top;
      }
      const unsigned p = bucket[top].back();
      bucket[top].pop_back();
      if (color[p] != UINT_MAX || nsat[p] != top) continue;
      const uint64_t *u = &used[(size_t)p * words];
      unsigned c = 0;
      for (size_t w = 0; w < words; ++w)
        if (~u[w]) {
          c = (unsigned)(64 * w) + (unsigned)__builtin_ctzll(~u[w]);
          break;
        }
      color[p] = c;
      nc = std::max(nc, c + 1);
      ++done;
      const size_t cw = c / 64;
      const uint64_t cb = 1ull << (c % 64);
      for (unsigned k = psup_ind[p]; k < psup_ind[p + 1]; ++k) {
        const unsigned q = psup[k];
        uint64_t &uq = used[(size_t)q * words + cw];
        if (color[q] != UINT_MAX || (uq & cb)) continue;
        uq |= cb;
        bucket[++nsat[q]].push_back(q);
        top = std::max<size_t>(top, nsat[q]);
      }
    }
    // nodes by colour, ascending old index inside a colour
    std::vector<unsigned> start(nc + 1, 0);
    for (size_t p = 0; p < np; ++p) start[color[p] + 1]++;
    for (unsigned c = 0; c < nc; ++c) start[c + 1] += start[c];
    for (size_t p = 0; p < np; ++p) new2old[start[color[p]]++] = (unsigned)p;
    *ncolor = nc;
  });
}

int dfm2h_tetmesh_create(dfm2h_tetmesh *out, const unsigned *tet, size_t ntet, size_t np) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(out, DFM2H_ERR_INVALID, "null output");
    DFM2H_REQUIRE(tet || ntet == 0, DFM2H_ERR_INVALID, "null tet");
    DFM2H_REQUIRE(ntet < (1ull << 32) && np < (1ull << 32), DFM2H_ERR_INVALID,
                  "mesh too large for 32-bit indices");
    for (size_t i = 0; i < ntet * 4; ++i)
      DFM2H_REQUIRE(tet[i] < np, DFM2H_ERR_INVALID, "tet node index out of range");
    auto *M = new dfm2h_tetmesh_s;
    try {
      M->serial = next_serial();
      M->ntet = ntet;
      M->np = np;
      std::vector<unsigned> ind(np + 1, 0), es(ntet * 4);
      for (size_t i = 0; i < ntet * 4; ++i) ind[tet[i] + 1]++;
      for (size_t p = 0; p < np; ++p) ind[p + 1] += ind[p];
      std::vector<unsigned> cur(ind.begin(), ind.end() - 1);
      for (size_t e = 0; e < ntet; ++e)
        for (int k = 0; k < 4; ++k) es[cur[tet[e * 4 + k]]++] = (unsigned)e;
      for (size_t q = 0; q < np; ++q) M->max_elsup = std::max(M->max_elsup, ind[q + 1] - ind[q]);
      // nodes beyond the LDS form of the device pattern build, and where
      // their candidate runs start (4 per element around the node)
      {
        std::vector<unsigned> big, coff;
        size_t tot = 0;
        for (size_t q = 0; q < np; ++q) {
          const unsigned d = ind[q + 1] - ind[q];
          if (d <= (unsigned)kPsupMaxElem) continue;
          big.push_back((unsigned)q);
          coff.push_back((unsigned)tot);
          tot += 4 * (size_t)d;
        }
        DFM2H_REQUIRE(tot < (1ull << 32), DFM2H_ERR_INVALID, "mesh too large for 32-bit indices");
        M->nbig = big.size();
        M->big_cand = tot;
        if (!big.empty()) {
          M->big.upload(big);
          M->big_off.upload(coff);
        }
      }
      M->tet.upload(tet, ntet * 4);
      M->elsup_ind.upload(ind);
      M->elsup.upload(es);
      DFM2H_HIP(hipStreamSynchronize(stream()));
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
  });
}

int dfm2h_psup_tet_device(dfm2h_tetmesh M, unsigned *psup_ind, unsigned *psup,
                          size_t psup_cap, size_t *npsup) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(M && psup_ind && npsup, DFM2H_ERR_INVALID, "null argument");
    const size_t np = M->np;
    psup_ind[0] = 0;
    if (np == 0) {
      *npsup = 0;
      return;
    }
    DBuf<unsigned> cnt, cand;
    cnt.alloc(np + 1);
    if (M->nbig) cand.alloc(M->big_cand);
    const dim3 g(ceil_div(np, kPsupBlock)), gb(ceil_div(std::max<size_t>(M->nbig, 1), kPsupBlock));
    auto rows = [&](const unsigned *ind, unsigned *out) {
      hipLaunchKernelGGL(k_psup_rows, g, dim3(kPsupBlock), 0, stream(), (unsigned)np, M->tet.p,
                         M->elsup_ind.p, M->elsup.p, ind, out);
      if (M->nbig)
        hipLaunchKernelGGL(k_psup_rows_big, gb, dim3(kPsupBlock), 0, stream(),
                           (unsigned)M->nbig, M->big.p, M->big_off.p, cand.p, M->tet.p,
                           M->elsup_ind.p, M->elsup.p, ind, out);
      check_launch("k_psup_rows");
    };
    rows(nullptr, cnt.p);
    DFM2H_HIP(hipMemcpyAsync(psup_ind + 1, cnt.p, np * sizeof(unsigned), hipMemcpyDeviceToHost,
                             stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
    size_t tot = 0;
    for (size_t q = 1; q <= np; ++q) {
      tot += psup_ind[q];
      DFM2H_REQUIRE(tot < (1ull << 32), DFM2H_ERR_INVALID, "pattern too large for 32-bit indices");
      psup_ind[q] = (unsigned)tot;
    }
    *npsup = tot;
    if (!psup || psup_cap < tot) return;
    DBuf<unsigned> ind, out;
    ind.upload(psup_ind, np + 1);
    out.alloc(std::max<size_t>(tot, 1));
    rows(ind.p, out.p);
    if (tot)
      DFM2H_HIP(hipMemcpyAsync(psup, out.p, tot * sizeof(unsigned), hipMemcpyDeviceToHost,
                               stream()));
    DFM2H_HIP(hipStreamSynchronize(stream()));
  });
}

int dfm2h_tetmesh_destroy(dfm2h_tetmesh M) {
  return guard(__func__, [&] { delete M; });
}

}  // extern "C"
