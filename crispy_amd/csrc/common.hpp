// Shared host/device definitions of the crispy_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "crispy_hip.h"

namespace crispy {

// Scoring constants of genDist (genMain.h:32-35) and the out-of-band value
// (genCPU.cpp:350,416,421).
constexpr int kGO = -10;
constexpr int kGE = -1;
constexpr int kMismatch = -4;
constexpr int kMatch = 5;
constexpr int kNeg = -1000000;

// Largest band half-width kB = (L1+L2)/(2 band) handled by the register
// resident NW kernel (one template instance per kB).  Full alignments
// (band <= 1), pairs with non-ACGT bytes and kB > kKbSliceMax run the
// generic kernel.
constexpr int kKbRegMax = 32;
// Wider bands up to kKbSliceMax run the lane-sliced kernel (W = 2kB+1 band
// cells split over up to 64 lanes of 24 cells each).
constexpr int kKbSliceMax = (64 * 24 - 1) / 2;
constexpr int kNumBuckets = kKbSliceMax + 2;  // [0, kKbSliceMax] + generic
constexpr int kGenericBucket = kKbSliceMax + 1;

// k-mer tile geometry.
constexpr int kTile = 128;        // reads per tile side
constexpr int kChunkWords = 16;   // bitmap words staged in LDS per step (24 for 40- and 48-word bitmaps)
constexpr int kBitmapSlack = 32;  // words after the last bitmap row: a last chunk may be staged whole
// Buckets of the hashed k = 6 candidate filter: 1,280, 1,536 or 2,048 (40,
// 48 or 64 bitmap words per read, in 8-word groups); fewer buckets mean less
// tile work and more candidates for the exact passes.
constexpr int kHashWordsMax = 64;
__host__ __device__ constexpr bool valid_hash_words(int w) { return w == 40 || w == 48 || w == 64; }
// The candidate tiles append to kCandSegs output segments, each with its own
// counter kSegStride u64 words (4 KB) from the next: one counter for the
// whole chip serialises the per-wave reservations (~600k atomics per
// 10^5-read run), separate counters on separate lines spread them.
constexpr int kCandSegs = 256;
constexpr int kSegStride = 512;

// Words of the k-mer presence bitmap of one read (4^K bits), padded to a
// multiple of kChunkWords.
__host__ __device__ constexpr int kmer_words(int K) {
  return ((1 << (2 * K)) / 32) < kChunkWords ? kChunkWords : ((1 << (2 * K)) / 32);
}

// In-band cells #{1<=i<=L2, 1<=j<=L1 : |i-j| <= kb}.
__host__ __device__ inline int64_t band_cells(int L1, int L2, int kb) {
  // row i holds columns [max(1, i-kb), min(L1, i+kb)], non-empty for i <= L1+kb
  int64_t total = 0;
  const int iend = L2 < L1 + kb ? L2 : L1 + kb;
  // piecewise linear in i with breakpoints at kb+1 and L1-kb
  int bps[4] = {1, kb + 1, L1 - kb, iend + 1};
  // sort the two inner breakpoints
  if (bps[1] > bps[2]) { int t = bps[1]; bps[1] = bps[2]; bps[2] = t; }
  int a = 1;
  for (int s = 1; s <= 3; ++s) {
    int b = bps[s];
    if (b > iend + 1) b = iend + 1;
    if (b > a) {
      // rows [a, b): evaluate the count at both ends; it is linear inside
      auto cnt = [&](int i) -> int64_t {
        int hi = i + kb < L1 ? i + kb : L1;
        int lo = i - kb > 1 ? i - kb : 1;
        return (int64_t)(hi - lo + 1);
      };
      const int64_t c0 = cnt(a), c1 = cnt(b - 1);
      total += (c0 + c1) * (int64_t)(b - a) / 2;
      a = b;
    }
  }
  return total;
}

}  // namespace crispy

// ------------------------------------------------------------------ host side
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

struct ReadStore {
  int32_t n = 0;
  int64_t nbases = 0;
  int32_t maxlen = 0;
  int32_t minlen = 0;
  int all_clean = 1;
  DevBuf bases, off, len, pack, poff, clean;
  std::vector<int32_t> h_len;
  std::vector<uint8_t> h_clean;
  int generation = 0;  // bumped on every upload (invalidates cached profiles)
};

// A captured graph of a launch sequence and the arguments it was captured
// with (pointers, sizes, scalars); captured again when they change.
struct GraphSegment {
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> key;
};

struct KmerState {
  // per-read profile for the current K
  int K = 0;
  int profile_gen = -1;
  DevBuf bitmap, T, rcnt, roff, rlist, excess;
  // k = 6 candidate filter: presence of the 4,096 codes hashed into
  // 32 x hash_words buckets, and T - (distinct buckets); `hashed` when the
  // profile holds them.  hash_buckets: the setting (0 = default).
  DevBuf hbitmap, hexcess;
  bool hashed = false;
  int hash_words = crispy::kHashWordsMax;
  int hash_buckets = 0;
  // the repeat lists back to back in read order (rlistc + roffc[r])
  DevBuf roffc, rlistc;
  // K = 7, 8: signed-sorted K-mer lists and the k = 6 candidate pairs
  DevBuf loff, lists, cand, cand_alt, cand_n;
  // candidate segments: counters (kCandSegs x kSegStride u64) and their
  // total and largest count (2 u64); cand holds kCandSegs x cand_seg_cap keys
  DevBuf seg, seg_stat;
  int kshift = 20;  // result keys are i << kshift | j
  int exact_pass = -1;  // K <= 6 exact counts: -1 auto (by pair count), 0 in the tile, 1 a pass of their own
  int64_t cand_seg_cap = 0;
  int64_t candidates = 0;
  // result
  DevBuf keys, keys_alt, dists, dists_alt, pairs, counter;
  int64_t cap = 0;
  int64_t count = 0;
  bool valid = false;
  GraphSegment g_profile, g_tile;  // captured profile pass and tile pass
};

struct NwState {
  DevBuf in_pairs, keys, keys_alt, idx, order, dist, hist, cells, flags, sel, nsel;
  DevBuf out_pairs, out_dists, scratch;
  DevBuf plan;  // NwPlan of the DP phase (bucket table, family task lists, counters)
  DevBuf stat;  // cells per family (generic last) summed over timed runs
  DevBuf csr_rows, csr_cols;  // CSR view of the (i, j)-ordered result
  DevBuf sort_k0, sort_k1, sort_b0, sort_b1, sort_pairs, sort_dists;  // (dist, i, j) sorts
  int slice_width = 0;  // sliced-kernel cells per lane: 0 = cost model, 24 or 32 = forced
  GraphSegment g_dp, g_filter;  // captured DP phase and filter
  int64_t count = 0;
  int64_t in_count = 0;  // input pairs of the last run (the filter's scan covers them)
  bool valid = false;
  // order of the pairs the result came from: the CSR view needs them sorted by
  // (i, j) with every i below its row count
  bool ij_sorted = false;
  int32_t max_row = -1;
};

// Input side (preprocess-equivalent ordering and merge).
struct PreState {
  DevBuf bases, off, freq, keys, keys_alt, flags, sel, nsel, csum, out_order, out_freq;
  DevBuf primer, scratch;  // primer filter: the primer, per-lane DP rows
};

struct KStat {
  double ms = 0;
  int64_t launches = 0;
  double units = 0;
};

struct PendingEvent {
  std::string name;
  hipEvent_t a, b;
  double units;
};

constexpr int kNumWorkers = 2;  // with the main stream, within the 4 hardware queues a process gets

struct crispy_ctx {
  int device = 0;
  int cus = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  // Side streams (default priority) for independent launches of one phase;
  // forked from and joined back into `stream` with events.
  hipStream_t workers[kNumWorkers] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kNumWorkers] = {};
  std::string err;
  ReadStore reads[2];
  KmerState km;
  NwState nw;
  PreState pre;
  DevBuf temp;  // rocPRIM temporary storage
  void *h_stage = nullptr;  // pinned host staging for small async uploads
  size_t h_stage_bytes = 0;
  hipEvent_t stage_ev = nullptr;  // the last upload from h_stage has been read
  void *h_hist = nullptr;  // pinned slot for small device read-backs
  bool timing = false;
  bool graphs = true;  // kernel chains through captured graphs (false: launched one by one)
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, KStat> stats;
  int64_t dev_bytes = 0;
  int64_t mem_limit = 0;  // device bytes the buffers may grow to (0: no limit)
};

namespace crispy {

int set_err(crispy_ctx *ctx, int code, const char *fmt, ...);
int ensure(crispy_ctx *ctx, DevBuf &b, size_t bytes);
void release(crispy_ctx *ctx, DevBuf &b);
// Grow-only pinned host staging buffer (ctx->h_stage) that the host may
// write: waits for the last upload from it (recorded in stage_ev) and
// synchronises the main stream before an old buffer is freed.
int ensure_host(crispy_ctx *ctx, size_t bytes);
void timer_begin(crispy_ctx *ctx, const char *name, hipEvent_t *a, hipStream_t s = nullptr);
void timer_end(crispy_ctx *ctx, const char *name, hipEvent_t a, double units, hipStream_t s = nullptr);
int fork_workers(crispy_ctx *ctx);
int join_workers(crispy_ctx *ctx);
// NW timers: *units = the cells behind an NW DP timer name (summed on the
// device while timing is on); other names leave *units alone.
int nw_timer_units(crispy_ctx *ctx, const char *name, double *units);
int nw_timer_reset(crispy_ctx *ctx);
void release_graphs(crispy_ctx *ctx);

}  // namespace crispy

#define CRISPY_HIP(ctx, call)                                                                \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return crispy::set_err((ctx), CRISPY_ERR_HIP, "%s failed: %s (%s:%d)", #call,         \
                             hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)

#define CRISPY_TRY(expr)     \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != CRISPY_OK)    \
      return rc_;            \
  } while (0)

#define CRISPY_LAUNCHED(ctx) CRISPY_HIP(ctx, hipGetLastError())
