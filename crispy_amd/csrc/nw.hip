// genDist on gfx950: banded / full Needleman-Wunsch distance per pair with
// the scoring, tie and end-cell rules of computeGenDist_CPU_band
// (genCPU.cpp:348-464) and computeGenDist_CPU_full (genCPU.cpp:132-225).
//
// Recurrence (s1 = read pair.x on the column axis j, s2 = read pair.y on the
// row axis i, kB = (L1+L2)/(2 band), cells with |i-j| <= kB):
//   diag = M(i-1,j-1) + (s2[i-1]==s1[j-1] ? 5 : -4)
//   up   = M(i,j-1) + pen_h(U(i,j-1))   (-1000000 when (i,j-1) is out of band)
//   left = M(i-1,j) + pen_v(U(i-1,j))   (-1000000 when (i-1,j) is out of band)
//   M = max(diag, up, left), ties resolved diag > up > left; U records it.
//   pen_h: diag -10, up -1, left 0, boundary 0;  pen_v: diag -10, left -1,
//   up 0, boundary 0.  ML/AL follow the chosen predecessor (+1 each; a
//   matching diagonal adds 0 to ML).  Boundary row 0 / column 0 cells
//   (within the band) are M 0, ML 0, AL = index.
//   End cell: the last maximum (>=) over the last row (i = L2) and over the
//   last column (j = L1); the last-row one wins only if strictly larger.
//   dist = (float)ML / (float)AL.
//
// Kernels:
//   nw_band_reg_kernel<KB>  one lane per pair, the whole band row of 2kB+1
//     cells held in VGPRs in band coordinates d = j - i + kB, updated in
//     place in ascending d (diag = old[d], left = old[d+1], up = new[d-1]).
//     Pairs are bucketed by kB (template) and sorted by R = min(L2, L1 + kB),
//     the pair's last row with an in-band cell, so the 64 lanes of a wave
//     walk rows in lock step and stop together.  Reads are 2-bit packed; a
//     row's match bits come from XOR of a sliding 2-bit window of read 1
//     with read 2's base replicated.  Requires ACGT-only reads.
//   nw_band_sliced: kKbRegMax < kB <= kKbSliceMax, G lanes per pair.
//   The register and sliced bodies run as wave tasks of a few fused launches
//   (one per occupancy family), not one launch per kB bucket.
//   nw_generic_kernel  one lane per pair, the row state in a lane-
//     interleaved global scratch, raw byte compares.  Any kB, full mode and
//     reads with other bytes (N, lowercase, ...).
#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.hpp"
#include "sort_config.hpp"
#include "graph_segment.hpp"
#include "nw_fused.hpp"

namespace crispy {

const int2 *kmer_result_pairs(crispy_ctx *ctx);

// Bucket keys (kB - kB_lo) << rbits | R with R < 2^rbits, kB_lo the smallest
// kB the reads allow; key_bits = kb_bits + rbits where 2^kb_bits exceeds the
// span of possible kB by at least one, so the all-ones key of the generic
// kernel's pairs sorts after every other (and as few radix passes as the
// read lengths need run: 2 instead of 3 for 150-550 bp reads at band 10).
struct BucketKey {
  int kb_lo, rbits, bits;
};

static BucketKey bucket_key(int kb_lo, int kb_top, int maxlen) {
  BucketKey k{kb_lo, 1, 0};
  while ((1 << k.rbits) <= std::min(maxlen, 2047)) ++k.rbits;
  int kb_bits = 0;
  while ((1 << kb_bits) < std::max(kb_top - kb_lo, 0) + 2) ++kb_bits;
  k.bits = kb_bits + k.rbits;
  return k;
}

// ---------------------------------------------------------------- bucketing
// Sort key per pair: kB << 11 | R, R = min(L1, L2) the DP's last row with an
// in-band cell, clamped to 2047, for the register and sliced kernels; all ones
// for the generic kernel.  The keys are radix-sorted with the pair indices as
// values; the sort is stable, so equal keys keep the (i, j) order of the
// input and a wave's pairs mostly share read i (one L1).  Per-bucket pair and
// cell counts are reduced in LDS first (a handful of buckets would serialise
// global atomics).
__global__ __launch_bounds__(256) void nw_bucket_kernel(
    const int2 *__restrict__ pairs, int64_t np, const int32_t *__restrict__ len,
    const uint8_t *__restrict__ clean, int band, BucketKey bk, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ idx, unsigned int *__restrict__ hist,
    unsigned long long *__restrict__ cells) {
  const uint32_t generic_key = (1u << bk.bits) - 1;
  __shared__ unsigned int s_hist[kNumBuckets];
  __shared__ unsigned long long s_cells[kNumBuckets];
  for (int b = threadIdx.x; b < kNumBuckets; b += blockDim.x) {
    s_hist[b] = 0;
    s_cells[b] = 0;
  }
  __syncthreads();
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int2 pr = pairs[p];
    const int L1 = len[pr.x], L2 = len[pr.y];
    uint32_t key = generic_key;
    int bucket = kGenericBucket;
    int64_t nc;
    if (band >= 2) {
      const int kb = (L1 + L2) / (2 * band);
      nc = band_cells(L1, L2, kb);
      if (kb <= kKbSliceMax && clean[pr.x] && clean[pr.y]) {
        // rows past the column read's length + kB hold no in-band cell, and
        // the DP runs its rows over the shorter read (nw_dp.hpp), so it ends
        // at row min(L1, L2)
        const int rows = min(L1, L2);
        key = ((uint32_t)(kb - bk.kb_lo) << bk.rbits) | (uint32_t)min(rows, 2047);
        bucket = kb;
      }
    } else {
      nc = (int64_t)L1 * L2;
    }
    keys[p] = key;
    idx[p] = (uint32_t)p;
    atomicAdd(&s_hist[bucket], 1u);
    atomicAdd(&s_cells[bucket], (unsigned long long)nc);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kNumBuckets; b += blockDim.x) {
    if (s_hist[b]) {
      atomicAdd(&hist[b], s_hist[b]);
      atomicAdd(&cells[b], s_cells[b]);
    }
  }
}

// --------------------------------------------------------- generic kernel
// Cell packing: x = M, y = U << 28 | ML << 14 | AL  (U: 0 boundary, 1 diag,
// 2 left, 3 up).
__device__ __forceinline__ int pen_h(uint32_t U) { return U == 1 ? kGO : (U == 3 ? kGE : 0); }
__device__ __forceinline__ int pen_v(uint32_t U) { return U == 1 ? kGO : (U == 2 ? kGE : 0); }
__device__ __forceinline__ uint32_t pk(uint32_t U, int ML, int AL) {
  return (U << 28) | ((uint32_t)ML << 14) | (uint32_t)AL;
}

__global__ __launch_bounds__(256) void nw_generic_kernel(
    const uint8_t *__restrict__ bases, const int64_t *__restrict__ off,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, int32_t count, int band, int2 *__restrict__ scratch,
    float *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int2 *S = scratch + g;
  for (int64_t t = g; t < count; t += stride) {
    const uint32_t pidx = order[t];
    const int2 pr = pairs[pidx];
    const int L1 = len[pr.x], L2 = len[pr.y];
    const uint8_t *s1 = bases + off[pr.x];
    const uint8_t *s2 = bases + off[pr.y];
    const bool full = band <= 1;
    const int kb = full ? L1 + L2 : (L1 + L2) / (2 * band);
    S[0] = make_int2(0, 0);
    const int r0 = min(L1, kb);
    for (int j = 1; j <= r0; ++j) S[j * stride] = make_int2(0, (int)pk(0, 0, j));
    int bestI = kNeg, bestJ = kNeg, MLI = 0, ALI = 0, MLJ = 0, ALJ = 0;
    for (int i = 1; i <= L2; ++i) {
      const int jlo = max(1, i - kb), jhi = min(L1, i + kb);
      if (jlo > jhi) break;
      const bool colb = i <= kb;  // boundary cell (i, 0) exists
      int2 prev_old = S[(int64_t)(jlo - 1) * stride];
      int2 up = make_int2(kNeg, 0);
      bool upv = false;
      if (jlo == 1 && colb) {
        up = make_int2(0, (int)pk(0, 0, i));
        S[0] = up;
        upv = true;
      }
      const uint8_t b = s2[i - 1];
      for (int j = jlo; j <= jhi; ++j) {
        const int2 old = S[(int64_t)j * stride];
        const bool leftv = j <= i - 1 + kb;
        const int mt = b == s1[j - 1];
        const int diag = prev_old.x + (mt ? kMatch : kMismatch);
        const uint32_t uu = (uint32_t)up.y, uo = (uint32_t)old.y, ud = (uint32_t)prev_old.y;
        const int upval = upv ? up.x + pen_h(uu >> 28) : kNeg;
        const int leftval = leftv ? old.x + pen_v(uo >> 28) : kNeg;
        const int M = max(max(diag, upval), leftval);
        uint32_t cell;
        if (M == diag)
          cell = pk(1, (int)((ud >> 14) & 0x3FFF) - mt + 1, (int)(ud & 0x3FFF) + 1);
        else if (M == upval)
          cell = pk(3, (int)((uu >> 14) & 0x3FFF) + 1, (int)(uu & 0x3FFF) + 1);
        else
          cell = pk(2, (int)((uo >> 14) & 0x3FFF) + 1, (int)(uo & 0x3FFF) + 1);
        const int2 nc = make_int2(M, (int)cell);
        S[(int64_t)j * stride] = nc;
        prev_old = old;
        up = nc;
        upv = true;
        if (i == L2 && M >= bestI) {
          bestI = M;
          MLI = (cell >> 14) & 0x3FFF;
          ALI = cell & 0x3FFF;
        }
        if (j == L1 && M >= bestJ) {
          bestJ = M;
          MLJ = (cell >> 14) & 0x3FFF;
          ALJ = cell & 0x3FFF;
        }
      }
    }
    out[pidx] = bestI > bestJ ? (float)MLI / (float)ALI : (float)MLJ / (float)ALJ;
  }
}

// ------------------------------------------------------------ filtering
__device__ __forceinline__ bool nw_passes(float dist, float thr) {
  const float diff = dist - thr;
  return dist < thr || (double)fabsf(diff) < 0.00001;
}

__global__ void nw_flag_kernel(const float *__restrict__ dist, int64_t np, float thr,
                               uint32_t *__restrict__ flags) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x)
    flags[p] = nw_passes(dist[p], thr) ? 1u : 0u;
}

// The passing pairs in input order: pos = exclusive scan of the flags; the
// count goes to *nsel.
__global__ void nw_compact_kernel(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ pos,
                                  int64_t np, const int2 *__restrict__ pairs, const float *__restrict__ dist,
                                  int2 *__restrict__ out_pairs, float *__restrict__ out_dists,
                                  unsigned int *__restrict__ nsel) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t f = flags[p], q = pos[p];
    if (f) {
      out_pairs[q] = pairs[p];
      out_dists[q] = dist[p];
    }
    if (p == np - 1) *nsel = q + f;
  }
}

__global__ void pack_pair_keys_kernel(const int2 *__restrict__ pairs, int64_t n,
                                      uint64_t *__restrict__ keys) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int2 p = pairs[k];
    keys[k] = ((uint64_t)(uint32_t)p.x << 20) | (uint64_t)(uint32_t)p.y;
  }
}

__global__ void dist_bits_kernel(const float *__restrict__ d, int64_t n, uint32_t *__restrict__ bits) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    bits[k] = __float_as_uint(d[k]);
}

__global__ void unpack_pair_keys_kernel(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ bits,
                                        int64_t n, int2 *__restrict__ pairs, float *__restrict__ d) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t v = keys[k];
    pairs[k] = make_int2((int)(v >> 20), (int)(v & 0xFFFFFu));
    d[k] = __uint_as_float(bits[k]);
  }
}

// CSR row starts of the (i, j)-ordered result: row_ptr[r] = first k with
// pairs[k].x >= r (lower bound), r = 0..n.
__global__ void nw_csr_rows_kernel(const int2 *__restrict__ pairs, int64_t count, int32_t n,
                                   int64_t *__restrict__ row_ptr) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n;
       r += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = count;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (pairs[mid].x < r) lo = mid + 1; else hi = mid;
    }
    row_ptr[r] = lo;
  }
}

__global__ void nw_csr_cols_kernel(const int2 *__restrict__ pairs, int64_t count, int32_t *__restrict__ cols) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x)
    cols[k] = pairs[k].y;
}

static inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }


using FamilyLaunch = void (*)(hipStream_t, unsigned, const NwFusedArgs &);
static const FamilyLaunch kFamilyLaunch[kNumFamilies] = {
    launch_nw_reg_family0, launch_nw_reg_family1, launch_nw_reg_family2, launch_nw_reg_family3,
    launch_nw_sliced_family24, launch_nw_sliced_family32};
static_assert(kNumFamilies == 6 && kSliceWidths[0] == 24 && kSliceWidths[1] == 32, "one launcher per family");

// ------------------------------------------------------------ DP phase plan
// One block of 1024 threads, thread b for bucket b: the plan of the DP phase
// from the bucket histogram (btab by an exclusive scan of the pair counts;
// per family its buckets by descending mean cells per pair, ties by bucket,
// and their first tasks).  With `stat` non-null the per-family cells (generic
// last) are added to it for the timers.
__global__ __launch_bounds__(1024) void nw_plan_kernel(const unsigned int *__restrict__ hist,
                                                       const unsigned long long *__restrict__ cells,
                                                       int forced, NwPlan *__restrict__ plan,
                                                       unsigned long long *__restrict__ stat) {
  static_assert(kNumBuckets <= 1024, "one thread per bucket");
  __shared__ uint32_t s_wsum[16];
  __shared__ int s_live[kNumBuckets];
  __shared__ int s_nlive;
  __shared__ double s_mean[kNumBuckets];
  __shared__ int s_fam[kNumBuckets];
  __shared__ uint32_t s_waves[kNumBuckets];
  __shared__ uint32_t s_ntasks[kNumFamilies], s_nbk[kNumFamilies];
  __shared__ unsigned long long s_cells[kNumFamilies + 1];
  const int b = threadIdx.x, lane = b & 63, wv = b >> 6;
  if (b == 0) s_nlive = 0;
  if (b < kNumFamilies) s_ntasks[b] = s_nbk[b] = 0;
  if (b <= kNumFamilies) s_cells[b] = 0;
  const uint32_t h = b < kNumBuckets ? hist[b] : 0u;
  // exclusive scan of h: within the wave, then over the 16 wave sums
  uint32_t x = h;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[wv] = x;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wv; ++w) before += s_wsum[w];
  const uint32_t start = before + x - h;
  int f = -1;
  uint32_t waves = 0;
  double mean = 0;
  if (b < kNumBuckets) {
    plan->btab[b] = make_uint2(start, h);
    if (b < kGenericBucket && h) {
      f = nw_family_of(b, forced);
      const uint32_t per = (uint32_t)nw_pairs_per_wave(b, f);
      waves = (h + per - 1) / per;
      mean = (double)cells[b] / (double)h;
      s_live[atomicAdd(&s_nlive, 1)] = b;
    }
    s_fam[b] = f;
    s_waves[b] = waves;
    s_mean[b] = mean;
  }
  __syncthreads();
  if (f >= 0) {
    uint32_t rank = 0, first = 0;
    for (int k = 0; k < s_nlive; ++k) {
      const int c = s_live[k];
      if (s_fam[c] == f && (s_mean[c] > mean || (s_mean[c] == mean && c < b))) {
        ++rank;
        first += s_waves[c];
      }
    }
    plan->list[f][rank] = make_uint2((uint32_t)b, first);
    atomicAdd(&s_ntasks[f], waves);
    atomicAdd(&s_nbk[f], 1u);
    atomicAdd(&s_cells[f], cells[b]);
  }
  if (b == kGenericBucket) s_cells[kNumFamilies] = cells[b];
  __syncthreads();
  if (b < kNumFamilies) {
    plan->ntasks[b] = s_ntasks[b];
    plan->nbuckets[b] = s_nbk[b];
  }
  if (stat && b <= kNumFamilies) stat[b] += s_cells[b];
}

// Grid of family f for np pairs whose kB lies in [kb_lo, kb_hi]: an upper
// bound on the family's tasks, np over the fewest pairs per wave among the
// family's kB in that range plus one partial wave per bucket, rounded up to
// m * 2^e with m < 16 (similar pair counts launch alike and share a captured
// graph; at most 1/8 of the blocks are spare).  0 if no kB of the range
// belongs to the family.
static int64_t family_grid(int f, int64_t np, int kb_lo, int kb_hi, int forced) {
  int64_t per = 0, buckets = 0;
  for (int kb = std::max(kb_lo, 0); kb <= std::min(kb_hi, kKbSliceMax); ++kb) {
    if (nw_family_of(kb, forced) != f) continue;
    const int64_t p = nw_pairs_per_wave(kb, f);
    per = per ? std::min(per, p) : p;
    ++buckets;
  }
  if (!buckets) return 0;
  int64_t g = (np + per - 1) / per + buckets;
  int e = 0;
  while ((g >> e) >= 16) ++e;
  return (((g - 1) >> e) + 1) << e;
}

// The DP phase's launches after the bucket sort: the plan, then every family
// that the reads' kB range allows on the worker streams (family f on worker
// f mod 2: neighbouring families, which common length spreads put to work
// together, run side by side).  No host read-back and nothing sized by the histogram, so the
// sequence can be captured into a graph.  Per-family timers only when not
// capturing.
static int enqueue_dp(crispy_ctx *ctx, const NwFusedArgs &args, const unsigned (&grid)[kNumFamilies],
                      bool capturing) {
  NwState &nw = ctx->nw;
  nw_plan_kernel<<<1, 1024, 0, ctx->stream>>>(nw.hist.as<unsigned int>(), nw.cells.as<unsigned long long>(),
                                              nw.slice_width, nw.plan.as<NwPlan>(),
                                              ctx->timing ? nw.stat.as<unsigned long long>() : nullptr);
  CRISPY_LAUNCHED(ctx);
  CRISPY_TRY(fork_workers(ctx));
  for (int f = 0; f < kNumFamilies; ++f) {
    if (!grid[f]) continue;
    hipStream_t s = ctx->workers[f % kNumWorkers];
    const char *name = f >= kNumRegFamilies ? "nw_band_sliced" : "nw_band_reg";
    hipEvent_t ev = nullptr;
    if (!capturing) timer_begin(ctx, name, &ev, s);
    kFamilyLaunch[f](s, grid[f], args);
    CRISPY_LAUNCHED(ctx);
    if (!capturing) timer_end(ctx, name, ev, 0.0, s);
  }
  return join_workers(ctx);
}

// Distances of np device pairs into nw.dist (indexed like the pairs): the
// bucketing (keys, radix sort), the DP phase (plan and family kernels; a
// captured graph with graphs on), then the generic kernel if the reads can
// have such pairs.
static int nw_core(crispy_ctx *ctx, const int2 *d_pairs, int64_t np, int band) {
  ReadStore &rs = ctx->reads[CRISPY_READS_NW];
  NwState &nw = ctx->nw;
  if (np > 0xFFFFFFFFll) return set_err(ctx, CRISPY_ERR_ARG, "too many pairs in one call");
  // kB = (L1 + L2) / (2 band) of any pair lies in [minlen / band, maxlen / band]
  const int kb_lo = band >= 2 ? rs.minlen / band : 0;
  const int kb_hi = band >= 2 ? rs.maxlen / band : -1;
  CRISPY_TRY(ensure(ctx, nw.keys, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.keys_alt, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.idx, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.order, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.dist, sizeof(float) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.hist, sizeof(unsigned int) * kNumBuckets));
  CRISPY_TRY(ensure(ctx, nw.cells, sizeof(unsigned long long) * kNumBuckets));
  CRISPY_TRY(ensure(ctx, nw.plan, sizeof(NwPlan)));
  if (ctx->timing && !nw.stat.p) {  // the cell counters start at zero (timing_reset clears them later)
    CRISPY_TRY(ensure(ctx, nw.stat, sizeof(unsigned long long) * (kNumFamilies + 1)));
    CRISPY_HIP(ctx, hipMemsetAsync(nw.stat.p, 0, nw.stat.bytes, ctx->stream));
  }
  if (np == 0) return CRISPY_OK;
  uint32_t *order = nw.order.as<uint32_t>();
  float *out = nw.dist.as<float>();
  NwPlan *plan = nw.plan.as<NwPlan>();

  // bucketing: launched one by one (the radix sort clears its own state
  // with memsets, which stay out of graphs)
  const BucketKey bkey = bucket_key(kb_lo, std::min(kb_hi, kKbSliceMax), rs.maxlen);
  hipEvent_t ev;
  CRISPY_HIP(ctx, hipMemsetAsync(nw.hist.p, 0, sizeof(unsigned int) * kNumBuckets, ctx->stream));
  CRISPY_HIP(ctx, hipMemsetAsync(nw.cells.p, 0, sizeof(unsigned long long) * kNumBuckets, ctx->stream));
  timer_begin(ctx, "nw_bucket", &ev);
  nw_bucket_kernel<<<(int)std::min<int64_t>((np + 255) / 256, 1024), 256, 0, ctx->stream>>>(
      d_pairs, np, rs.len.as<int32_t>(), rs.clean.as<uint8_t>(), band, bkey, nw.keys.as<uint32_t>(),
      nw.idx.as<uint32_t>(), nw.hist.as<unsigned int>(), nw.cells.as<unsigned long long>());
  CRISPY_LAUNCHED(ctx);
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(nullptr, tbytes, nw.keys.as<uint32_t>(),
                                            nw.keys_alt.as<uint32_t>(), nw.idx.as<uint32_t>(), order,
                                            (size_t)np, 0, bkey.bits, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(ctx->temp.p, tbytes, nw.keys.as<uint32_t>(),
                                            nw.keys_alt.as<uint32_t>(), nw.idx.as<uint32_t>(), order,
                                            (size_t)np, 0, bkey.bits, ctx->stream));
  timer_end(ctx, "nw_bucket", ev, (double)np);

  const NwFusedArgs args{rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(),
                         d_pairs, order, plan, out};
  unsigned grid[kNumFamilies];
  for (int f = 0; f < kNumFamilies; ++f) {
    const int64_t g = family_grid(f, np, kb_lo, kb_hi, nw.slice_width);
    if (g > 0x7FFFFFFF)
      return set_err(ctx, CRISPY_ERR_ARG, "%lld pairs need %lld wave tasks in one launch; split the call",
                     (long long)np, (long long)g);
    grid[f] = (unsigned)g;
  }
  std::vector<uintptr_t> key = {
      (uintptr_t)args.pack, (uintptr_t)args.poff, (uintptr_t)args.len, (uintptr_t)args.pairs,
      (uintptr_t)args.order, (uintptr_t)args.plan, (uintptr_t)args.out, (uintptr_t)nw.hist.p,
      (uintptr_t)nw.cells.p, (uintptr_t)(ctx->timing ? nw.stat.p : nullptr), (uintptr_t)nw.slice_width};
  for (unsigned g : grid) key.push_back(g);
  hipEvent_t phase;
  timer_begin(ctx, "nw_dp_phase", &phase);
  CRISPY_TRY(run_segment(ctx, nw.g_dp, std::move(key), "NW DP phase",
                         [&](bool capturing) { return enqueue_dp(ctx, args, grid, capturing); }));
  // Pairs the fused kernels cannot take (full NW, non-ACGT bytes, kB past
  // kKbSliceMax) exist only if the reads allow them; then their count comes
  // back from the plan and the generic kernel runs after the families.
  const bool maybe_generic = band < 2 || !rs.all_clean || rs.maxlen / band > kKbSliceMax;
  if (maybe_generic) {
    uint2 *h_g = reinterpret_cast<uint2 *>(ctx->h_hist);
    CRISPY_HIP(ctx, hipMemcpyAsync(h_g, &plan->btab[kGenericBucket], sizeof(uint2), hipMemcpyDeviceToHost,
                                   ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint2 gb = *h_g;
    const int64_t ng = gb.y;
    if (ng > 0) {
      const int gblocks = (int)((std::min<int64_t>(ng, 32768) + 255) / 256);
      CRISPY_TRY(ensure(ctx, nw.scratch, sizeof(int2) * (size_t)gblocks * 256 * (rs.maxlen + 1)));
      timer_begin(ctx, "nw_generic", &ev);
      nw_generic_kernel<<<gblocks, 256, 0, ctx->stream>>>(rs.bases.as<uint8_t>(), rs.off.as<int64_t>(),
                                                         rs.len.as<int32_t>(), d_pairs, order + gb.x,
                                                         (int32_t)ng, band, nw.scratch.as<int2>(), out);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "nw_generic", ev, 0.0);
    }
  }
  timer_end(ctx, "nw_dp_phase", phase, 0.0);
  return CRISPY_OK;
}

// Keeps the pairs whose distance passes, in input order (flags, their
// exclusive scan, a compaction; a captured graph with graphs on).
static int nw_filter(crispy_ctx *ctx, const int2 *d_pairs, int64_t np, float thr, int64_t *count) {
  NwState &nw = ctx->nw;
  CRISPY_TRY(ensure(ctx, nw.flags, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.sel, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.nsel, sizeof(unsigned int)));
  CRISPY_TRY(ensure(ctx, nw.out_pairs, sizeof(int2) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.out_dists, sizeof(float) * (np + 1)));
  *count = 0;
  if (np == 0) return CRISPY_OK;
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::exclusive_scan(nullptr, tbytes, nw.flags.as<uint32_t>(), nw.sel.as<uint32_t>(), 0u,
                                          (size_t)np, rocprim::plus<uint32_t>(), ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  uint32_t thr_bits;
  std::memcpy(&thr_bits, &thr, sizeof thr_bits);
  hipEvent_t ev;
  timer_begin(ctx, "nw_filter", &ev);
  CRISPY_TRY(run_segment(
      ctx, nw.g_filter,
      {(uintptr_t)d_pairs, (uintptr_t)np, thr_bits, (uintptr_t)nw.dist.p, (uintptr_t)nw.flags.p,
       (uintptr_t)nw.sel.p, (uintptr_t)nw.nsel.p, (uintptr_t)nw.out_pairs.p, (uintptr_t)nw.out_dists.p,
       (uintptr_t)ctx->temp.p, (uintptr_t)tbytes},
      "NW filter", [&](bool) -> int {
        nw_flag_kernel<<<grid_for(np), 256, 0, ctx->stream>>>(nw.dist.as<float>(), np, thr, nw.flags.as<uint32_t>());
        CRISPY_LAUNCHED(ctx);
        size_t tb = tbytes;
        CRISPY_HIP(ctx, rocprim::exclusive_scan(ctx->temp.p, tb, nw.flags.as<uint32_t>(), nw.sel.as<uint32_t>(),
                                                0u, (size_t)np, rocprim::plus<uint32_t>(), ctx->stream));
        nw_compact_kernel<<<grid_for(np), 256, 0, ctx->stream>>>(
            nw.flags.as<uint32_t>(), nw.sel.as<uint32_t>(), np, d_pairs, nw.dist.as<float>(),
            nw.out_pairs.as<int2>(), nw.out_dists.as<float>(), nw.nsel.as<unsigned int>());
        CRISPY_LAUNCHED(ctx);
        return CRISPY_OK;
      }));
  timer_end(ctx, "nw_filter", ev, (double)np);
  unsigned int *h_c = reinterpret_cast<unsigned int *>(ctx->h_hist);
  CRISPY_HIP(ctx, hipMemcpyAsync(h_c, nw.nsel.p, sizeof *h_c, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *count = *h_c;
  return CRISPY_OK;
}

static int upload_pairs(crispy_ctx *ctx, const int32_t *pairs, int64_t np) {
  const int32_t n = ctx->reads[CRISPY_READS_NW].n;
  for (int64_t p = 0; p < 2 * np; ++p)
    if (pairs[p] < 0 || pairs[p] >= n)
      return set_err(ctx, CRISPY_ERR_ARG, "pair %lld names read %d outside [0, %d)", (long long)(p / 2),
                     pairs[p], n);
  bool sorted = true;
  int32_t max_row = -1;
  for (int64_t p = 0; p < np; ++p) {
    max_row = std::max(max_row, pairs[2 * p]);
    if (p > 0 && (pairs[2 * p] < pairs[2 * p - 2] ||
                  (pairs[2 * p] == pairs[2 * p - 2] && pairs[2 * p + 1] < pairs[2 * p - 1])))
      sorted = false;
  }
  ctx->nw.ij_sorted = sorted;
  ctx->nw.max_row = max_row;
  CRISPY_TRY(ensure(ctx, ctx->nw.in_pairs, sizeof(int2) * (np + 1)));
  if (np > 0)
    CRISPY_HIP(ctx, hipMemcpyAsync(ctx->nw.in_pairs.p, pairs, sizeof(int2) * np, hipMemcpyHostToDevice,
                                   ctx->stream));
  return CRISPY_OK;
}

int nw_timer_units(crispy_ctx *ctx, const char *name, double *units) {
  NwState &nw = ctx->nw;
  int lo, hi;  // families [lo, hi); kNumFamilies is the generic kernel
  if (!std::strcmp(name, "nw_band_reg")) lo = 0, hi = kNumRegFamilies;
  else if (!std::strcmp(name, "nw_band_sliced")) lo = kNumRegFamilies, hi = kNumFamilies;
  else if (!std::strcmp(name, "nw_generic")) lo = kNumFamilies, hi = kNumFamilies + 1;
  else if (!std::strcmp(name, "nw_dp_phase")) lo = 0, hi = kNumFamilies + 1;
  else return CRISPY_OK;
  *units = 0;
  if (!nw.stat.p) return CRISPY_OK;
  unsigned long long *h = reinterpret_cast<unsigned long long *>(ctx->h_hist);
  CRISPY_HIP(ctx, hipMemcpyAsync(h, nw.stat.p, sizeof(unsigned long long) * (kNumFamilies + 1),
                                 hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int f = lo; f < hi; ++f) *units += (double)h[f];
  return CRISPY_OK;
}

int nw_timer_reset(crispy_ctx *ctx) {
  if (!ctx->nw.stat.p) return CRISPY_OK;
  CRISPY_HIP(ctx, hipMemsetAsync(ctx->nw.stat.p, 0, ctx->nw.stat.bytes, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

// Sorts n (pair, dist) entries by (dist, i, j) into out_pairs / out_dists
// (device buffers that may alias the inputs).  Distances are >= 0, so their
// bit patterns order like the values.  Input already in (i, j) order needs
// one stable sort by distance; other input is first sorted by (i, j).
static int sort_dist_ij(crispy_ctx *ctx, const int2 *pairs, const float *dists, int64_t n, bool ij_sorted,
                        int2 *out_pairs, float *out_dists) {
  NwState &nw = ctx->nw;
  CRISPY_TRY(ensure(ctx, nw.sort_k0, sizeof(uint64_t) * n));
  CRISPY_TRY(ensure(ctx, nw.sort_k1, sizeof(uint64_t) * n));
  CRISPY_TRY(ensure(ctx, nw.sort_b0, sizeof(uint32_t) * n));
  CRISPY_TRY(ensure(ctx, nw.sort_b1, sizeof(uint32_t) * n));
  uint64_t *k0 = nw.sort_k0.as<uint64_t>(), *k1 = nw.sort_k1.as<uint64_t>();
  uint32_t *b0 = nw.sort_b0.as<uint32_t>(), *b1 = nw.sort_b1.as<uint32_t>();
  pack_pair_keys_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(pairs, n, k0);
  dist_bits_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dists, n, b0);
  CRISPY_LAUNCHED(ctx);
  size_t t1 = 0, t2 = 0;
  if (!ij_sorted)
    CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(nullptr, t1, k0, k1, b0, b1, (size_t)n, 0, 40,
                                                             ctx->stream));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(nullptr, t2, b0, b1, k0, k1, (size_t)n, 0, 32,
                                                           ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, std::max(t1, t2)));
  if (!ij_sorted) {
    // 1) by (i, j) into k1/b1, 2) stable by dist back into k0/b0
    CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(ctx->temp.p, t1, k0, k1, b0, b1, (size_t)n, 0, 40,
                                                             ctx->stream));
    CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(ctx->temp.p, t2, b1, b0, k1, k0, (size_t)n, 0, 32,
                                                             ctx->stream));
    unpack_pair_keys_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(k0, b0, n, out_pairs, out_dists);
  } else {
    CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(ctx->temp.p, t2, b0, b1, k0, k1, (size_t)n, 0, 32,
                                                             ctx->stream));
    unpack_pair_keys_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(k1, b1, n, out_pairs, out_dists);
  }
  CRISPY_LAUNCHED(ctx);
  return CRISPY_OK;
}

// ------------------------------------------------------ no-allocation form
// crispy_nw_dists_noalloc keeps everything beyond its inputs and outputs in
// `scratch`, one array per device that lives as long as the process.  In host
// code the array's name is not a device address: its address on a device
// comes from hipGetSymbolAddress with that device current (noalloc_scratch).
// Layout, each part 256-byte aligned: bucket keys, sorted keys, pair indices,
// sorted indices (4 bytes per pair each), the bucket histogram and cell
// counts, the DP plan, a bad-pair flag, then the radix sort's temporary
// storage.
constexpr int64_t kNoAllocMaxPairs = int64_t(1) << 20;
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kNaKeys = 0;
constexpr size_t kNaKeysAlt = kNaKeys + 4 * kNoAllocMaxPairs;
constexpr size_t kNaIdx = kNaKeysAlt + 4 * kNoAllocMaxPairs;
constexpr size_t kNaOrder = kNaIdx + 4 * kNoAllocMaxPairs;
constexpr size_t kNaHist = kNaOrder + 4 * kNoAllocMaxPairs;
constexpr size_t kNaCells = kNaHist + align256(sizeof(unsigned int) * kNumBuckets);
constexpr size_t kNaPlan = kNaCells + align256(sizeof(unsigned long long) * kNumBuckets);
constexpr size_t kNaFlag = kNaPlan + align256(sizeof(NwPlan));
constexpr size_t kNaTemp = kNaFlag + 256;
constexpr size_t kNaTempBytes = size_t(24) << 20;
constexpr size_t kNoAllocScratchBytes = kNaTemp + kNaTempBytes;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Sets *bad when a pair names a read outside [0, n).
__global__ void nw_check_pairs_kernel(const int2 *__restrict__ pairs, int64_t np, int32_t n,
                                      unsigned int *__restrict__ bad) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int2 q = pairs[p];
    if (q.x < 0 || q.x >= n || q.y < 0 || q.y >= n) *bad = 1u;
  }
}

// Device address of `scratch` on `device`, looked up once per device with the
// device current (the caller has made it current).
static int noalloc_scratch(crispy_ctx *ctx, unsigned char **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static unsigned char *addr[kMaxDevices] = {};
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_err(ctx, CRISPY_ERR_ARG, "device %d outside [0, %d)", ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[ctx->device]) {
    void *ptr = nullptr;
    CRISPY_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    addr[ctx->device] = static_cast<unsigned char *>(ptr);
  }
  *out = addr[ctx->device];
  return CRISPY_OK;
}

void release_graphs(crispy_ctx *ctx) {
  drop_segment(ctx->nw.g_dp);
  drop_segment(ctx->nw.g_filter);
  drop_segment(ctx->km.g_profile);
  drop_segment(ctx->km.g_tile);
}

}  // namespace crispy

using namespace crispy;

extern "C" {

int crispy_set_graphs(crispy_ctx *ctx, int enable) {
  if (!ctx) return CRISPY_ERR_ARG;
  ctx->graphs = enable != 0;
  return CRISPY_OK;
}

int crispy_nw_run(crispy_ctx *ctx, const int32_t *pairs, int64_t npairs, int band, float threshold,
                  int64_t *count) {
  if (!ctx || npairs < 0 || (npairs > 0 && !pairs)) return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  ctx->nw.valid = false;
  CRISPY_TRY(upload_pairs(ctx, pairs, npairs));
  const int2 *d = ctx->nw.in_pairs.as<int2>();
  CRISPY_TRY(nw_core(ctx, d, npairs, band));
  int64_t c = 0;
  CRISPY_TRY(nw_filter(ctx, d, npairs, threshold, &c));
  ctx->nw.count = c;
  ctx->nw.in_count = npairs;
  ctx->nw.valid = true;
  if (count) *count = c;
  return CRISPY_OK;
}

int crispy_nw_run_on_kmer(crispy_ctx *ctx, int band, float threshold, int64_t *count) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (!ctx->km.valid) return set_err(ctx, CRISPY_ERR_STATE, "no k-mer result to consume");
  if (ctx->reads[CRISPY_READS_NW].n != ctx->reads[CRISPY_READS_KMER].n)
    return set_err(ctx, CRISPY_ERR_STATE, "NW read slot does not hold the k-mer reads");
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  ctx->nw.valid = false;
  const int64_t np = ctx->km.count;
  const int2 *d = kmer_result_pairs(ctx);
  // the k-mer result is sorted by (i, j) and names reads of this slot
  ctx->nw.ij_sorted = true;
  ctx->nw.max_row = ctx->reads[CRISPY_READS_NW].n - 1;
  CRISPY_TRY(nw_core(ctx, d, np, band));
  int64_t c = 0;
  CRISPY_TRY(nw_filter(ctx, d, np, threshold, &c));
  ctx->nw.count = c;
  ctx->nw.in_count = np;
  ctx->nw.valid = true;
  if (count) *count = c;
  return CRISPY_OK;
}

int crispy_nw_fetch(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs, float *dists) {
  if (!ctx) return CRISPY_ERR_ARG;
  NwState &nw = ctx->nw;
  if (!nw.valid) return set_err(ctx, CRISPY_ERR_STATE, "no NW result");
  if (first < 0 || count < 0 || first + count > nw.count)
    return set_err(ctx, CRISPY_ERR_ARG, "fetch range outside the result");
  if (count == 0) return CRISPY_OK;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  if (pairs)
    CRISPY_HIP(ctx, hipMemcpyAsync(pairs, nw.out_pairs.as<int2>() + first, sizeof(int2) * count,
                                   hipMemcpyDeviceToHost, ctx->stream));
  if (dists)
    CRISPY_HIP(ctx, hipMemcpyAsync(dists, nw.out_dists.as<float>() + first, sizeof(float) * count,
                                   hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int crispy_nw_file_split(crispy_ctx *ctx, int64_t chunk_pairs, int64_t flush_pairs, int64_t *file_ends,
                         int32_t max_files, int32_t *nfiles) {
  if (!ctx || chunk_pairs < 1 || flush_pairs < 1 || max_files < 0 || !nfiles || (max_files > 0 && !file_ends))
    return CRISPY_ERR_ARG;
  NwState &nw = ctx->nw;
  if (!nw.valid) return set_err(ctx, CRISPY_ERR_STATE, "no NW result");
  const int64_t np = nw.in_count;
  const int64_t chunks = (np + chunk_pairs - 1) / chunk_pairs;
  // the filter's exclusive scan: sel[p] = passing pairs among inputs [0, p)
  std::vector<uint32_t> ends((size_t)chunks, 0);
  if (chunks > 0) {
    CRISPY_HIP(ctx, hipSetDevice(ctx->device));
    for (int64_t c = 0; c + 1 < chunks; ++c)
      CRISPY_HIP(ctx, hipMemcpyAsync(&ends[(size_t)c], nw.sel.as<uint32_t>() + (c + 1) * chunk_pairs,
                                     sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ends[(size_t)chunks - 1] = (uint32_t)nw.count;
  }
  int32_t f = 0;
  int64_t begin = 0;
  for (int64_t c = 0; c < chunks; ++c) {
    if ((int64_t)ends[(size_t)c] - begin >= flush_pairs) {
      if (f < max_files) file_ends[f] = ends[(size_t)c];
      ++f;
      begin = ends[(size_t)c];
    }
  }
  if (nw.count > begin) {
    if (f < max_files) file_ends[f] = nw.count;
    ++f;
  }
  *nfiles = f;
  return f <= max_files ? CRISPY_OK : set_err(ctx, CRISPY_ERR_ARG, "%d files, room for %d", f, max_files);
}

int crispy_nw_fetch_sorted(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs, float *dists) {
  if (!ctx || (count > 0 && (!pairs || !dists))) return CRISPY_ERR_ARG;
  NwState &nw = ctx->nw;
  if (!nw.valid) return set_err(ctx, CRISPY_ERR_STATE, "no NW result");
  if (first < 0 || count < 0 || first + count > nw.count)
    return set_err(ctx, CRISPY_ERR_ARG, "fetch range outside the result");
  if (count == 0) return CRISPY_OK;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  CRISPY_TRY(ensure(ctx, nw.sort_pairs, sizeof(int2) * count));
  CRISPY_TRY(ensure(ctx, nw.sort_dists, sizeof(float) * count));
  CRISPY_TRY(sort_dist_ij(ctx, nw.out_pairs.as<int2>() + first, nw.out_dists.as<float>() + first, count,
                          nw.ij_sorted, nw.sort_pairs.as<int2>(), nw.sort_dists.as<float>()));
  CRISPY_HIP(ctx, hipMemcpyAsync(pairs, nw.sort_pairs.p, sizeof(int2) * count, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(dists, nw.sort_dists.p, sizeof(float) * count, hipMemcpyDeviceToHost,
                                 ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int crispy_nw_autotune(crispy_ctx *ctx, int band, int repeats, int *chosen) {
  if (!ctx || repeats < 1) return CRISPY_ERR_ARG;
  if (!ctx->km.valid) return set_err(ctx, CRISPY_ERR_STATE, "no k-mer result to tune on");
  if (ctx->reads[CRISPY_READS_NW].n != ctx->reads[CRISPY_READS_KMER].n)
    return set_err(ctx, CRISPY_ERR_STATE, "NW read slot does not hold the k-mer reads");
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  NwState &nw = ctx->nw;
  const int64_t np = ctx->km.count;
  const int2 *d = kmer_result_pairs(ctx);
  const int policies[] = {0, kSliceWidths[0], kSliceWidths[1]};
  double best = 1e300;
  int pick = nw.slice_width;
  for (int pol : policies) {
    nw.slice_width = pol;
    double fastest = 1e300;
    for (int r = 0; r <= repeats; ++r) {  // the first run warms up
      CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const auto t0 = std::chrono::steady_clock::now();
      CRISPY_TRY(nw_core(ctx, d, np, band));
      CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (r > 0) fastest = std::min(fastest, s);
    }
    if (fastest < best) {
      best = fastest;
      pick = pol;
    }
  }
  nw.slice_width = pick;
  if (chosen) *chosen = pick;
  return CRISPY_OK;
}

int crispy_nw_set_slice_width(crispy_ctx *ctx, int cells) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (cells != 0 && cells != kSliceWidths[0] && cells != kSliceWidths[1])
    return set_err(ctx, CRISPY_ERR_ARG, "slice width %d: 0 (cost model), %d or %d", cells, kSliceWidths[0],
                   kSliceWidths[1]);
  ctx->nw.slice_width = cells;
  return CRISPY_OK;
}

int crispy_nw_csr(crispy_ctx *ctx, int32_t n_rows, int64_t *row_ptr, int32_t *cols, float *vals) {
  if (!ctx || n_rows < 0 || !row_ptr) return CRISPY_ERR_ARG;
  NwState &nw = ctx->nw;
  if (!nw.valid) return set_err(ctx, CRISPY_ERR_STATE, "no NW result");
  if (!nw.ij_sorted)
    return set_err(ctx, CRISPY_ERR_STATE, "NW result is not in (i, j) order: its input pairs were not sorted");
  if (nw.max_row >= n_rows)
    return set_err(ctx, CRISPY_ERR_ARG, "NW result names row %d, the CSR has %d rows", nw.max_row, n_rows);
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t c = nw.count;
  CRISPY_TRY(ensure(ctx, nw.csr_rows, sizeof(int64_t) * ((size_t)n_rows + 1)));
  CRISPY_TRY(ensure(ctx, nw.csr_cols, sizeof(int32_t) * (c + 1)));
  nw_csr_rows_kernel<<<grid_for((int64_t)n_rows + 1), 256, 0, ctx->stream>>>(
      nw.out_pairs.as<int2>(), c, n_rows, nw.csr_rows.as<int64_t>());
  CRISPY_LAUNCHED(ctx);
  if (c > 0) {
    nw_csr_cols_kernel<<<grid_for(c), 256, 0, ctx->stream>>>(nw.out_pairs.as<int2>(), c,
                                                           nw.csr_cols.as<int32_t>());
    CRISPY_LAUNCHED(ctx);
  }
  CRISPY_HIP(ctx, hipMemcpyAsync(row_ptr, nw.csr_rows.p, sizeof(int64_t) * ((size_t)n_rows + 1),
                                 hipMemcpyDeviceToHost, ctx->stream));
  if (c > 0 && cols)
    CRISPY_HIP(ctx, hipMemcpyAsync(cols, nw.csr_cols.p, sizeof(int32_t) * c, hipMemcpyDeviceToHost, ctx->stream));
  if (c > 0 && vals)
    CRISPY_HIP(ctx, hipMemcpyAsync(vals, nw.out_dists.p, sizeof(float) * c, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int crispy_nw_dists_noalloc(crispy_ctx *ctx, const int32_t *d_pairs, int64_t npairs, int band,
                            float *d_dists) {
  if (!ctx || npairs < 0 || (npairs > 0 && (!d_pairs || !d_dists))) return CRISPY_ERR_ARG;
  if (npairs > kNoAllocMaxPairs)
    return set_err(ctx, CRISPY_ERR_ARG, "%lld pairs: the no-allocation form takes at most %lld per call",
                   (long long)npairs, (long long)kNoAllocMaxPairs);
  ReadStore &rs = ctx->reads[CRISPY_READS_NW];
  if (rs.n <= 0) return set_err(ctx, CRISPY_ERR_STATE, "no reads in the NW slot");
  // only the register and lane-sliced kernels: no per-lane DP scratch
  if (band < 2 || !rs.all_clean || rs.maxlen / band > kKbSliceMax)
    return set_err(ctx, CRISPY_ERR_ARG,
                   "the no-allocation form needs band >= 2, ACGT-only reads and kB <= %d (band %d, "
                   "longest read %d)", kKbSliceMax, band, rs.maxlen);
  if (npairs == 0) return CRISPY_OK;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  unsigned char *base = nullptr;
  CRISPY_TRY(noalloc_scratch(ctx, &base));
  uint32_t *keys = reinterpret_cast<uint32_t *>(base + kNaKeys);
  uint32_t *keys_alt = reinterpret_cast<uint32_t *>(base + kNaKeysAlt);
  uint32_t *idx = reinterpret_cast<uint32_t *>(base + kNaIdx);
  uint32_t *order = reinterpret_cast<uint32_t *>(base + kNaOrder);
  unsigned int *hist = reinterpret_cast<unsigned int *>(base + kNaHist);
  unsigned long long *cells = reinterpret_cast<unsigned long long *>(base + kNaCells);
  NwPlan *plan = reinterpret_cast<NwPlan *>(base + kNaPlan);
  unsigned int *bad = reinterpret_cast<unsigned int *>(base + kNaFlag);
  void *temp = base + kNaTemp;
  const int2 *pairs = reinterpret_cast<const int2 *>(d_pairs);
  const int kb_lo = rs.minlen / band, kb_hi = rs.maxlen / band;
  const BucketKey bkey = bucket_key(kb_lo, std::min(kb_hi, kKbSliceMax), rs.maxlen);
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(nullptr, tbytes, keys, keys_alt, idx, order,
                                                           (size_t)npairs, 0, bkey.bits, ctx->stream));
  if (tbytes > kNaTempBytes)
    return set_err(ctx, CRISPY_ERR_STATE, "radix sort needs %zu bytes of scratch, %zu reserved", tbytes,
                   kNaTempBytes);
  // the previous call's contents stay in the array: clear what starts at zero
  CRISPY_HIP(ctx, hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream));
  CRISPY_HIP(ctx, hipMemsetAsync(hist, 0, sizeof(unsigned int) * kNumBuckets, ctx->stream));
  CRISPY_HIP(ctx, hipMemsetAsync(cells, 0, sizeof(unsigned long long) * kNumBuckets, ctx->stream));
  nw_check_pairs_kernel<<<grid_for(npairs), 256, 0, ctx->stream>>>(pairs, npairs, rs.n, bad);
  CRISPY_LAUNCHED(ctx);
  unsigned int *h_bad = reinterpret_cast<unsigned int *>(ctx->h_hist);
  CRISPY_HIP(ctx, hipMemcpyAsync(h_bad, bad, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (*h_bad) return set_err(ctx, CRISPY_ERR_ARG, "a pair names a read outside [0, %d)", rs.n);
  nw_bucket_kernel<<<(int)std::min<int64_t>((npairs + 255) / 256, 1024), 256, 0, ctx->stream>>>(
      pairs, npairs, rs.len.as<int32_t>(), rs.clean.as<uint8_t>(), band, bkey, keys, idx, hist, cells);
  CRISPY_LAUNCHED(ctx);
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(temp, tbytes, keys, keys_alt, idx, order,
                                                           (size_t)npairs, 0, bkey.bits, ctx->stream));
  const NwFusedArgs args{rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(),
                         pairs, order, plan, d_dists};
  nw_plan_kernel<<<1, 1024, 0, ctx->stream>>>(hist, cells, ctx->nw.slice_width, plan, nullptr);
  CRISPY_LAUNCHED(ctx);
  CRISPY_TRY(fork_workers(ctx));
  for (int f = 0; f < kNumFamilies; ++f) {
    const int64_t g = family_grid(f, npairs, kb_lo, kb_hi, ctx->nw.slice_width);
    if (!g) continue;
    kFamilyLaunch[f](ctx->workers[f % kNumWorkers], (unsigned)g, args);
    CRISPY_LAUNCHED(ctx);
  }
  CRISPY_TRY(join_workers(ctx));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int crispy_nw_dists(crispy_ctx *ctx, const int32_t *pairs, int64_t npairs, int band, float *dists) {
  if (!ctx || npairs < 0 || (npairs > 0 && (!pairs || !dists))) return CRISPY_ERR_ARG;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  ctx->nw.valid = false;
  CRISPY_TRY(upload_pairs(ctx, pairs, npairs));
  CRISPY_TRY(nw_core(ctx, ctx->nw.in_pairs.as<int2>(), npairs, band));
  if (npairs > 0)
    CRISPY_HIP(ctx, hipMemcpyAsync(dists, ctx->nw.dist.p, sizeof(float) * npairs, hipMemcpyDeviceToHost,
                                   ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

int crispy_sort_by_dist(crispy_ctx *ctx, int32_t *pairs, float *dists, int64_t n) {
  if (!ctx || n < 0 || (n > 0 && (!pairs || !dists))) return CRISPY_ERR_ARG;
  if (n <= 1) return CRISPY_OK;
  for (int64_t k = 0; k < 2 * n; ++k)
    if (pairs[k] < 0 || pairs[k] >= CRISPY_MAX_NUM_READS)
      return set_err(ctx, CRISPY_ERR_ARG, "read id %d outside [0, 2^20)", pairs[k]);
  for (int64_t k = 0; k < n; ++k)
    if (!(dists[k] >= 0.0f))
      return set_err(ctx, CRISPY_ERR_ARG, "distance %g is not >= 0", (double)dists[k]);
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  NwState &nw = ctx->nw;
  CRISPY_TRY(ensure(ctx, nw.sort_pairs, sizeof(int2) * n));
  CRISPY_TRY(ensure(ctx, nw.sort_dists, sizeof(float) * n));
  CRISPY_HIP(ctx, hipMemcpyAsync(nw.sort_pairs.p, pairs, sizeof(int2) * n, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(nw.sort_dists.p, dists, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_TRY(sort_dist_ij(ctx, nw.sort_pairs.as<int2>(), nw.sort_dists.as<float>(), n, false,
                          nw.sort_pairs.as<int2>(), nw.sort_dists.as<float>()));
  CRISPY_HIP(ctx, hipMemcpyAsync(pairs, nw.sort_pairs.p, sizeof(int2) * n, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(dists, nw.sort_dists.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

}  // extern "C"
