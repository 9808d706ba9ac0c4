// kmerDist on gfx950.
//
// The reference merges two sorted uint16 k-mer lists per pair
// (kmerCPU.cpp:29-43, kmerKernel.cu:37-45).  That merge counts
//     matches = I_hi + [hinfo_a == hinfo_b] * I_lo
// where I_x = sum_v min(c_a[v], c_b[v]) over the codes of part x, "hi" are
// the codes >= 0x8000 (only possible for K = 8, first base T or C), and
// hinfo = (largest hi code, its multiplicity) or 0: the lists are sorted as
// signed short (kmerMain.cpp:156-159, 275-279) and compared unsigned, so the
// lo part is reached only when both hi parts run out together.  For K <= 7
// every code is lo and hinfo is 0, so matches is the multiset intersection.
//
// sum_v min(c_a, c_b) = sum_v [c_a>=1][c_b>=1] + sum_v min(c_a-1, c_b-1)^+,
// so for K <= 6 each read is stored as
//   * a presence bitmap of 4^K bits (padded to 16-word chunks), and
//   * a short sorted list of (code << 16 | count-1) for codes seen twice or
//     more ("repeats"),
// and a 128x128 tile of pairs is a {0,1} matrix product of the bitmaps on the
// matrix cores plus a merge of the two repeat lists for the pairs the
// popcount bound does not settle.  The filter dist <= thr and
// the compaction of passing pairs run in the same kernel; only passing pairs
// leave the chip.  K = 7, 8 use the k = 6 tiles as a candidate filter and the
// reference's merge on the candidates (kmer_cand_kernel).
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <utility>
#include <vector>

#include "common.hpp"
#include "graph_segment.hpp"
#include "sort_config.hpp"

namespace crispy {

// --------------------------------------------------------------- profiles
// Bucket of a k = 6 code in the hashed candidate filter of 32 x HW buckets:
// the multiplicative hash h = code * 2654435761 mod 2^32 scaled to the bucket
// range, (h * buckets) >> 32 (for 2,048 buckets its top 11 bits).  Merging
// codes into buckets can only raise sum_b min(count_a(b), count_b(b)), so the
// filter's bound stays an upper bound of the k-mer matches.
constexpr int kDefaultHashWords = 40;  // 1,280 buckets: the tile work saved outweighs the extra candidates (reads of 150-900 bp)
template <int HW>
__host__ __device__ inline uint32_t kmer_bucket(uint32_t code) {
  static_assert(valid_hash_words(HW), "1,280, 1,536 or 2,048 buckets");
  return (uint32_t)(((uint64_t)(code * 2654435761u) * (uint64_t)(32 * HW)) >> 32);
}

// One block of 128 threads per read: the read's packed words are staged in
// LDS, the codes of its T = L-K+1 k-mers (first base most significant, as
// READ::formTuples kmerMain.cpp:117-154) counted in an LDS histogram of 16-bit
// bins (two per word: a count is at most T <= 2,048) over all 4^K codes; the
// bins then give the presence bitmap, the repeat entries in code order and
// excess = T - (distinct codes).  HW > 0 (K = 6) also writes the presence of
// the codes' buckets (kmer_bucket<HW>, HW words per read) and
// hexcess = T - (distinct buckets).  (Small blocks and half-size bins: 16
// reads in flight per CU, against 8 with 256 threads and 32-bit bins.)
template <int K, int HW = 0>
__global__ __launch_bounds__(128) void kmer_profile_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, uint32_t *__restrict__ bitmap, int32_t *__restrict__ Tout,
    int32_t *__restrict__ rcnt, const int64_t *__restrict__ roff, uint32_t *__restrict__ rlist,
    int32_t *__restrict__ excess, uint32_t *__restrict__ hbitmap, int32_t *__restrict__ hexcess) {
  constexpr int NT = 128;
  constexpr int WP = kmer_words(K);
  constexpr int NBINS = 1 << (2 * K);
  constexpr int NHW = NBINS >= 2 ? NBINS / 2 : 1;       // histogram words (two 16-bit bins each)
  constexpr int BPT = NBINS >= NT ? NBINS / NT : 1;     // bins per thread (within one bitmap word)
  constexpr int kMaxWords = (CRISPY_MAX_READ_LEN + 15) / 16;
  constexpr bool HASH = HW > 0;
  static_assert(!HASH || K == 6, "the hashed filter covers k = 6");
  static_assert(BPT <= 32 && (BPT == 1 || BPT % 2 == 0), "a thread's bins lie in one bitmap word");
  __shared__ uint32_t hist[NHW];
  __shared__ uint32_t sread[kMaxWords + 1];
  __shared__ uint32_t bm[WP];
  __shared__ uint32_t hbm[HASH ? HW : 1];
  __shared__ int wsum[2];
  __shared__ int wrep[NHW >= NT ? NHW / NT : 1][2];  // repeat entries per (step, wave)
  __shared__ int s_distinct, s_hdistinct;
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const int L = len[r];
  const int T = L - K + 1;  // >= 1, checked by the host
  const int nwr = (L + 15) >> 4;
  const uint32_t *s = pack + poff[r];
  for (int w = tid; w < nwr; w += NT) sread[w] = s[w];
  if (tid == 0) sread[nwr] = 0;  // the last window's high word
  for (int v = tid; v < NHW; v += NT) hist[v] = 0;
  if (NHW < NT || NBINS / 32 < WP)  // words the transposed scan does not store (K = 4 pads to 16)
    for (int w = tid; w < WP; w += NT) bm[w] = 0;
  if (HASH)
    for (int w = tid; w < HW; w += NT) hbm[w] = 0;
  if (tid == 0) s_distinct = s_hdistinct = 0;
  __syncthreads();
  for (int t = tid; t < T; t += NT) {
    // bases t .. t+K-1 (2 bits each, base p at bits 2(p mod 16) of word p / 16)
    const uint32_t win = __builtin_amdgcn_alignbit(sread[(t >> 4) + 1], sread[t >> 4], 2 * (t & 15));
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) c = (c << 2) | ((win >> (2 * q)) & 3u);
    atomicAdd(&hist[c >> 1], 1u << (16 * (c & 1)));
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  if constexpr (NHW >= NT) {
    // Transposed scan: at step k thread tid reads histogram word v = k * NT + tid
    // (bins 2v and 2v + 1), so a wave's reads hit distinct banks.  Ballots over
    // the even and odd bins give the wave's 128 presence bits: bitmap word
    // w = 8k + 4wv + g holds lane 16g + i's even bin (code 32w + 2i) at bit i
    // and its odd bin (code 32w + 2i + 1) at bit 16 + i, stored by lane g with
    // no atomics.  (For K >= 4 the bits of a word are in this order, the same
    // for every read; the passes that read the bitmaps take popcounts of two
    // reads' words ANDed, which no fixed order within a word changes.)
    constexpr int NSTEP = NHW / NT;
    static_assert(NT == 128 && NSTEP * 8 <= WP, "two waves of 64 lanes, 2 bins each, per step");
    int ndist = 0;  // wave-uniform
    for (int k = 0; k < NSTEP; ++k) {
      const int v = k * NT + tid;
      const uint32_t hw = hist[v], c0 = hw & 0xFFFFu, c1 = hw >> 16;
      const uint64_t e = __ballot(c0 > 0), o = __ballot(c1 > 0);
      const uint64_t r0 = __ballot(c0 >= 2), r1 = __ballot(c1 >= 2);
      ndist += __popcll(e) + __popcll(o);
      if (lane < 4) bm[8 * k + 4 * wv + lane] = ((uint32_t)(e >> (16 * lane)) & 0xFFFFu) | ((uint32_t)(o >> (16 * lane)) << 16);
      if (lane == 0) wrep[k][wv] = __popcll(r0) + __popcll(r1);
      if (HASH) {
        if (c0) {
          const uint32_t h = kmer_bucket<HASH ? HW : kHashWordsMax>((uint32_t)(2 * v));
          atomicOr(&hbm[h >> 5], 1u << (h & 31));
        }
        if (c1) {
          const uint32_t h = kmer_bucket<HASH ? HW : kHashWordsMax>((uint32_t)(2 * v + 1));
          atomicOr(&hbm[h >> 5], 1u << (h & 31));
        }
      }
    }
    if (lane == 0) wsum[wv] = ndist;
    __syncthreads();
    // The repeat entries in code order: step k's after every earlier step's,
    // wave 0's before wave 1's, and within a wave by lane, even bin first.
    uint32_t *out = rlist + roff[r];
    int run = 0;  // entries of the steps before k
    for (int k = 0; k < NSTEP; ++k) {
      const int n0 = wrep[k][0], n1 = wrep[k][1];
      if (wv ? n1 : n0) {  // wave-uniform
        const int v = k * NT + tid;
        const uint32_t hw = hist[v], c0 = hw & 0xFFFFu, c1 = hw >> 16;
        const bool q0 = c0 >= 2, q1 = c1 >= 2;
        const uint64_t below = (1ull << lane) - 1;
        int pos = run + (wv ? n0 : 0) + __popcll(__ballot(q0) & below) + __popcll(__ballot(q1) & below);
        if (q0) out[pos++] = ((uint32_t)(2 * v) << 16) | (c0 - 1);
        if (q1) out[pos] = ((uint32_t)(2 * v + 1) << 16) | (c1 - 1);
      }
      run += n0 + n1;
    }
    if (tid == 0) {
      rcnt[r] = run;
      Tout[r] = T;
      excess[r] = T - wsum[0] - wsum[1];  // sum over codes of (count - 1)
    }
  } else {
    // thread tid owns bins [tid * BPT, tid * BPT + BPT)
    const int b0 = tid * BPT;
    auto bin = [&](int v) -> uint32_t { return (hist[v >> 1] >> (16 * (v & 1))) & 0xFFFFu; };
    uint32_t bits = 0;
    int nrep = 0, ndist = 0;
    if (b0 < NBINS) {
#pragma unroll
      for (int k = 0; k < BPT; ++k) {
        const uint32_t c = bin(b0 + k);
        bits |= (c > 0 ? 1u : 0u) << k;
        ndist += c > 0;
        nrep += c >= 2;
        if (HASH && c > 0) {
          const uint32_t h = kmer_bucket<HASH ? HW : kHashWordsMax>((uint32_t)(b0 + k));
          atomicOr(&hbm[h >> 5], 1u << (h & 31));
        }
      }
      if (bits) atomicOr(&bm[b0 >> 5], bits << (b0 & 31));
    }
    if (ndist) atomicAdd(&s_distinct, ndist);
    // where this thread's repeat entries go: an inclusive scan of nrep within
    // the wave, then the total of the wave before it
    int incl = nrep;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    const int before = wv > 0 ? wsum[0] : 0;
    if (nrep) {
      uint32_t *out = rlist + roff[r] + (before + incl - nrep);
#pragma unroll
      for (int k = 0; k < BPT; ++k) {
        const uint32_t c = bin(b0 + k);
        if (c >= 2) *out++ = ((uint32_t)(b0 + k) << 16) | (c - 1);
      }
    }
    if (tid == NT - 1) rcnt[r] = before + incl;  // the last thread's inclusive sum: every entry
    if (tid == 0) {
      Tout[r] = T;
      excess[r] = T - s_distinct;  // sum over codes of (count - 1)
    }
  }
  uint32_t *dst = bitmap + (int64_t)r * WP;
  for (int w = tid; w < WP; w += NT) dst[w] = bm[w];
  if (HASH) {  // the barrier after the scan ordered the bucket bits before these reads
    if (tid < HW) {
      const uint32_t w = hbm[tid];
      hbitmap[(int64_t)r * HW + tid] = w;
      atomicAdd(&s_hdistinct, __popc(w));
    }
    __syncthreads();
    if (tid == 0) hexcess[r] = T - s_hdistinct;
  }
}

// sum_v min(e_a[v], e_b[v]) over two repeat lists sorted by code (entries
// code << 16 | count - 1).
template <class P>
__device__ __forceinline__ uint32_t repeat_min_sum(P A, int na, P B, int nb) {
  uint32_t sum = 0;
  int x = 0, y = 0;
  while (x < na && y < nb) {
    const uint32_t ea = A[x], eb = B[y];
    const uint32_t va = ea >> 16, vb = eb >> 16;
    if (va == vb) sum += min(ea & 0xFFFFu, eb & 0xFFFFu);
    x += va <= vb;
    y += va >= vb;
  }
  return sum;
}

// The same sum over two lists in global memory, read W entries of each per
// memory round trip.  Every entry of a window whose code is at most the
// smaller of the two windows' last codes is consumed (so one window is used up
// whole each time), and a common code always lies in both windows at once
// (codes are unique within a list), so comparing all pairs of the two windows
// finds each exactly once.  Past a list's end a window holds code 0xFFFF (A)
// or 0xFFFE (B): these match nothing.
template <int W>
__device__ __forceinline__ uint32_t repeat_min_sum_windowed(const uint32_t *__restrict__ A, int na,
                                                            const uint32_t *__restrict__ B, int nb) {
  uint32_t sum = 0;
  int x = 0, y = 0;
  while (x < na && y < nb) {
    uint32_t a[W], b[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
      a[s] = x + s < na ? A[x + s] : 0xFFFFFFFFu;
      b[s] = y + s < nb ? B[y + s] : 0xFFFEFFFFu;
    }
    const uint32_t lim = min(a[W - 1], b[W - 1]) >> 16;
    int da = 0, db = 0;
#pragma unroll
    for (int s = 0; s < W; ++s) {
      da += (a[s] >> 16) <= lim;
      db += (b[s] >> 16) <= lim;
#pragma unroll
      for (int t = 0; t < W; ++t)
        if ((a[s] ^ b[t]) < 0x10000u) sum += min(a[s] & 0xFFFFu, b[t] & 0xFFFFu);
    }
    x += da;
    y += db;
  }
  return sum;
}

// Candidate entries of the tile pass: the pair key i << kshift | j (at most
// 40 bits) and, above it, the pair's presence popcount (at most 4096).
constexpr int kCandPopShift = 44;
// Codes per list in a lane's merge window of the K = 7, 8 candidate kernel,
// and the codes past the last K-mer list that a window may read.
constexpr int kCandWin = 32;
constexpr int kListSlack = 32;
static_assert(kCandWin % 4 == 0 && kCandWin <= kListSlack, "a window starts on a word at or before the list's last code");
// Pairs from which the K <= 6 exact counts run as a pass of their own (auto mode).
constexpr int64_t kExactPassPairs = 1000000000;
constexpr uint64_t kCandKeyMask = (1ull << 40) - 1;

// kmerMain.cpp:180 / kmerCPU.cpp:45
__device__ __forceinline__ bool passes(float dist, float thr) {
  const float diff = dist - thr;
  return dist < thr || (double)fabsf(diff) < 0.00001;
}

__host__ __device__ inline void tile_coords(int64_t t, int nb, int &bi, int &bj) {
  // t = bi*nb - bi*(bi-1)/2 + (bj - bi), bi <= bj < nb
  const double b = 2.0 * nb + 1.0;
  int r = (int)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);
  r = max(0, min(r, nb - 1));
  auto start = [nb](int row) -> int64_t { return (int64_t)row * nb - (int64_t)row * (row - 1) / 2; };
  while (r > 0 && start(r) > t) --r;
  while (r + 1 < nb && start(r + 1) <= t) ++r;
  bi = r;
  bj = r + (int)(t - start(r));
}

// ------------------------------------------------------ matrix-core tiles
// Tile order: the upper-triangular grid of 128x128 tiles is cut into
// kSuperTiles x kSuperTiles super-tiles, enumerated row-major over I <= J; each
// super-tile owns kSuperTiles^2 slots, its tiles row-major.  Slots below the
// diagonal or past the last tile are empty (a few per cent of the slots), so
// a slot's tile follows from its index alone.  Each XCD takes a contiguous run
// of slots, so the tiles it has in flight share a few dozen row and column
// blocks, which stay in its L2 (a plain row-major order re-streams every
// column block once per tile row).
constexpr int kSuperTiles = 16;
constexpr int kSuperSlots = kSuperTiles * kSuperTiles;

// tile (bi, bj) of slot t for nb tile rows; false for an empty slot
__host__ __device__ inline bool slot_coords(int64_t t, int nb, int &bi, int &bj) {
  const int ns = (nb + kSuperTiles - 1) / kSuperTiles;
  int I, J;
  tile_coords(t / kSuperSlots, ns, I, J);
  const int o = (int)(t % kSuperSlots);
  bi = I * kSuperTiles + o / kSuperTiles;
  bj = J * kSuperTiles + o % kSuperTiles;
  return bi <= bj && bj < nb;
}

__host__ __device__ inline int64_t num_slots(int nb) {
  const int64_t ns = (nb + kSuperTiles - 1) / kSuperTiles;
  return ns * (ns + 1) / 2 * kSuperSlots;
}

// One 128x128 tile of read pairs per block.  The popcount product runs on the
// matrix cores: popcount(a AND b) = sum_k a_k * b_k over {0,1} codes, as
// v_mfma_scale_f32_32x32x64_f8f6f4 with both operands FP4 (E2M1).  Operand
// lanes (checked on the device): lane l holds row (A) / column (B) l % 32 and
// 32 k-slots, the same 32 for A and B; accumulator r of lane l is pair
// (8(r/4) + 4(l/32) + r%4, l % 32).  Which 32 of a step's 64 codes a lane holds
// is free as long as A and B agree, so for an 8-word group lane half g takes
// words 4g..4g+3 and step s takes bit s of every nibble: w & (0x11111111 << s)
// for s = 0, 1, 2 (one VALU op per word; the nibbles read 0.5, 1 and 2) and
// (w >> 2) & 0x22222222 for s = 3 (bit 3 is the FP4 sign; the nibbles read 1).
// The A scale 2^2, 2^0, 2^-2, 2^0 makes every product 1, so the fp32
// accumulator holds the popcount exactly (at most 4096).
// Block: 2 waves, wave w owns tile rows [64w, 64w + 64) x all 128 columns
// (2 x 4 fragments, 128 accumulators per lane).  Bitmap chunks of 16 words are
// staged row-major (padded) in LDS, the next chunk prefetched into registers;
// tiles are mapped XCD-contiguously (work-group b -> XCD b mod 8 gets a
// contiguous run, so tiles sharing a row block share an L2).
// Epilogue: matches <= popcount + min(excess_i, excess_j) (excess = sum of
// count-1 over a read's codes), so a pair with
//   popcount + min(excess_i, excess_j) < min(F_i, F_j),  F = (T - tshift) * cut - 1
// cannot pass (F rises with T, so min(F_i, F_j) = F(min(T_i, T_j)); the margin
// in cut covers the fp32 roundings).  The tile tests the slightly weaker
//   popcount < min(G_i, G_j),  G = F - excess
// (if F_i <= F_j: min(F) - min(E) = F_i - min(E) >= F_i - E_i >= min(G); the
// same with i, j swapped), which needs one per-read value: 3 VALU ops per pair
// (the minimum of the non-negative G as the unsigned minimum of its bit
// patterns, the difference, its sign shifted into a 16-bit reject mask per
// fragment).  G <= 0 is stored as 0: such a pair survives either way.  A
// positive F - E is exact in fp32 (E is an integer below F).
// Diagonal tiles (i < j) and the tiles of the last row or column block (reads
// past n) mask their fragments explicitly.  The few survivors of the wave's 8
// fragments are queued in LDS and processed 64 at a time for the exact count
// (popcount plus the repeat-list merge when both reads have repeats) and the
// threshold; one atomic per flush (normally one per wave) reserves the
// output.  When the block has survivors, the tile's repeat lists are copied
// into the then idle chunk buffer (two coalesced runs of the compacted lists)
// if they fit, so the merges read LDS.
// CAND (K = 7, 8 candidates, tshift = K - 6): no merge, the bound alone.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), unrolled
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

typedef int mfma_v8i __attribute__((ext_vector_type(8)));
typedef float mfma_v16f __attribute__((ext_vector_type(16)));
typedef float lds_f4 __attribute__((ext_vector_type(4)));
constexpr int kFp4 = 4;          // f8f6f4 operand format code of FP4 (E2M1)
constexpr int kUnitScale = 127;  // E8M0 scale 2^0

// min of two non-negative floats (no NaN) as the unsigned min of their bits
__device__ __forceinline__ float umin_f(float a, float b) {
  return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int wave_incl_sum(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

template <int WP, bool CAND, int CW = kChunkWords>
__global__ __launch_bounds__(128, 2) void kmer_tile_mfma_kernel(
    const uint32_t *__restrict__ bitmap, const int32_t *__restrict__ Tn,
    const int32_t *__restrict__ rcnt, const int64_t *__restrict__ roff,
    const uint32_t *__restrict__ rlist, const int32_t *__restrict__ excess, int32_t n, int nb,
    int64_t tile_begin, int64_t ntiles, int64_t per_xcd, float thr, float cut, int tshift, int kshift,
    unsigned long long *__restrict__ counter, uint64_t *__restrict__ out_keys,
    float *__restrict__ out_dists, int64_t cap) {
  static_assert(WP % 8 == 0 && WP <= 128, "bitmap tiles hold at most 4,096 presence bits, in 8-word groups");
  // chunks of CW words (16, or 24 for 40- and 48-word bitmaps: 2 chunks, 2
  // barrier pairs per tile); the last chunk may hold fewer 8-word groups (the
  // rest of it, read past the row, is staged but not used: the bitmap buffers
  // end with kBitmapSlack words of slack)
  static_assert(CW == 16 || CW == 24, "16- or 24-word chunks");
  constexpr int NCH = (WP + CW - 1) / CW;
  constexpr int kLastGroups = (WP - (NCH - 1) * CW) / 8;  // 8-word groups in the last chunk
  constexpr int kRow = CW + 4;  // padded row stride of a staged chunk (words)
  constexpr int kStageWords = 2 * kTile * kRow;
  __shared__ uint32_t Ps[kStageWords];  // bitmap chunk in the main loop, then repeat lists
  __shared__ __align__(16) float sG[2 * kTile];
  __shared__ int32_t sT[2 * kTile];
  __shared__ int32_t sRc[2 * kTile];
  __shared__ int64_t sRo[2 * kTile];
  __shared__ int32_t sLo[2 * kTile];
  __shared__ int32_t sWave[2];
  __shared__ uint32_t sQ[2][1024];  // per-wave survivor queue

  const int64_t bid = blockIdx.x;
  if (CAND) {  // this block's output segment and its counter; cap is a segment's room
    const int seg = (int)(bid & (kCandSegs - 1));
    counter += (size_t)seg * kSegStride;
    out_keys += (size_t)seg * cap;
  }
  const int64_t local = (bid & 7) * per_xcd + (bid >> 3);
  if (local >= ntiles) return;
  int bi, bj;
  if (!slot_coords(tile_begin + local, nb, bi, bj)) return;  // block-uniform
  const int row0 = bi * kTile, col0 = bj * kTile;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int g = lane >> 5, r32 = lane & 31;

  // loader: a chunk is CW words of each of the 256 reads.  TPR lanes share a
  // read, each loading UPT 16-byte pieces of it (16-word chunks: lane quads,
  // one piece each; 24-word chunks: lane pairs, three each), so a wave
  // instruction reads whole runs; the block's 128 lanes cover RP reads of the
  // rows and of the columns per pass, NPASS passes.  Rows past n load read 0
  // (always valid): their pairs are masked in the epilogue.  Chunk 0 is
  // requested first and the per-read data right after it, so their
  // latencies overlap.
  constexpr int TPR = CW == 16 ? 4 : 2, UPT = CW / 4 / TPR;
  constexpr int RP = 128 / TPR, NPASS = kTile / RP;
  constexpr int NLD = NPASS * UPT;  // 16-byte pieces per lane per side
  const int lq = tid % TPR, lr = tid / TPR;
  uint32_t oa[NPASS], ob[NPASS];  // word offsets of this lane's first piece, chunk 0
#pragma unroll
  for (int k = 0; k < NPASS; ++k) {
    const int ra = row0 + lr + RP * k, rb = col0 + lr + RP * k;
    oa[k] = (uint32_t)(ra < n ? ra : 0) * WP + 4 * UPT * lq;
    ob[k] = (uint32_t)(rb < n ? rb : 0) * WP + 4 * UPT * lq;
  }
  // (built field by field: a whole-struct copy into the arrays keeps them in scratch)
  auto ld = [bitmap](uint32_t o) {
    const uint4 v = *reinterpret_cast<const uint4 *>(bitmap + o);
    return make_uint4(v.x, v.y, v.z, v.w);
  };
  uint4 pa[NLD], pb[NLD];
#pragma unroll
  for (int k = 0; k < NPASS; ++k)
#pragma unroll
    for (int u = 0; u < UPT; ++u) {
      pa[k * UPT + u] = ld(oa[k] + 4 * u);
      pb[k * UPT + u] = ld(ob[k] + 4 * u);
    }

  // per read of the tile (thread t: reads t and t + 128 of the 256)
  static_assert(2 * kTile == 2 * 128, "128 threads cover the tile's 256 reads in two steps");
  int pT[2], pE[2], pC[2];
  int64_t pO[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = tid + 128 * h;
    const int r = q < kTile ? row0 + q : col0 + (q - kTile);
    const int rr = r < n ? r : 0;  // unconditional loads (read 0 stands in), so none waits on another
    pT[h] = Tn[rr];
    pE[h] = excess[rr];
    if (!CAND) {  // the candidate form needs only the bound
      pC[h] = r < n ? rcnt[rr] : 0;
      pO[h] = r < n ? roff[rr] : 0;
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = tid + 128 * h;
    const int r = q < kTile ? row0 + q : col0 + (q - kTile);
    if (!CAND) {
      sT[q] = r < n ? pT[h] : 1;
      sRc[q] = pC[h];
      sRo[q] = pO[h];
    }
    // G = F - E (exact in fp32 when positive: E is an integer below F), >= 0;
    // 0 past the last read (those pairs are masked in the epilogue)
    const float G = fmaf((float)(pT[h] - tshift), cut, -1.0f) - (float)pE[h];
    sG[q] = r < n && G > 0.0f ? G : 0.0f;
  }

  mfma_v16f acc[2][4];
#pragma unroll
  for (int fa = 0; fa < 2; ++fa)
#pragma unroll
    for (int fb = 0; fb < 4; ++fb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[fa][fb][r] = 0.0f;

  auto mfma_step = [&](const uint4 *xa, const uint4 *xb, uint32_t M, int shift, int scale) {
    mfma_v8i ea[2], eb[4];
#pragma unroll
    for (int fa = 0; fa < 2; ++fa) {
      ea[fa][0] = (int)((xa[fa].x >> shift) & M);
      ea[fa][1] = (int)((xa[fa].y >> shift) & M);
      ea[fa][2] = (int)((xa[fa].z >> shift) & M);
      ea[fa][3] = (int)((xa[fa].w >> shift) & M);
      ea[fa][4] = ea[fa][5] = ea[fa][6] = ea[fa][7] = 0;  // unused by FP4 operands
    }
#pragma unroll
    for (int fb = 0; fb < 4; ++fb) {
      eb[fb][0] = (int)((xb[fb].x >> shift) & M);
      eb[fb][1] = (int)((xb[fb].y >> shift) & M);
      eb[fb][2] = (int)((xb[fb].z >> shift) & M);
      eb[fb][3] = (int)((xb[fb].w >> shift) & M);
      eb[fb][4] = eb[fb][5] = eb[fb][6] = eb[fb][7] = 0;
    }
#pragma unroll
    for (int fa = 0; fa < 2; ++fa)
#pragma unroll
      for (int fb = 0; fb < 4; ++fb)
        acc[fa][fb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ea[fa], eb[fb], acc[fa][fb], kFp4, kFp4, 0,
                                                                      scale, 0, kUnitScale);
  };

  for (int ch = 0; ch < NCH; ++ch) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NPASS; ++k)
#pragma unroll
      for (int u = 0; u < UPT; ++u) {
        *reinterpret_cast<uint4 *>(&Ps[(lr + RP * k) * kRow + 4 * (UPT * lq + u)]) = pa[k * UPT + u];
        *reinterpret_cast<uint4 *>(&Ps[(kTile + lr + RP * k) * kRow + 4 * (UPT * lq + u)]) = pb[k * UPT + u];
      }
    __syncthreads();
    if (ch + 1 < NCH) {  // prefetch the next chunk while this one is consumed
#pragma unroll
      for (int k = 0; k < NPASS; ++k)
#pragma unroll
        for (int u = 0; u < UPT; ++u) {
          pa[k * UPT + u] = ld(oa[k] + (ch + 1) * CW + 4 * u);
          pb[k * UPT + u] = ld(ob[k] + (ch + 1) * CW + 4 * u);
        }
    }
#pragma unroll
    for (int m = 0; m < CW / 8; ++m) {  // 8-word groups
      if (ch == NCH - 1 && m >= kLastGroups) break;  // block-uniform
      uint4 xa[2], xb[4];
#pragma unroll
      for (int fa = 0; fa < 2; ++fa)
        xa[fa] = *reinterpret_cast<const uint4 *>(&Ps[(wv * 64 + fa * 32 + r32) * kRow + 8 * m + 4 * g]);
#pragma unroll
      for (int fb = 0; fb < 4; ++fb)
        xb[fb] = *reinterpret_cast<const uint4 *>(&Ps[(kTile + fb * 32 + r32) * kRow + 8 * m + 4 * g]);
      if constexpr (CAND) {
      // the 4 bit planes software-pipelined in the wave: a column operand is
      // re-extracted for the next plane right after the two MFMAs that read it,
      // the row operands into a second set, so the wave's own VALU work runs
      // between its MFMAs instead of in front of them (tile -6 % at 2.5e5
      // reads; pinning the order with sched_group_barrier measured slower; the
      // exact-count form's larger epilogue state spills with it)
      auto ext = [](const uint4 &x, int pl) {
        const uint32_t M = pl < 3 ? 0x11111111u << pl : 0x22222222u;
        const int sh = pl < 3 ? 0 : 2;
        mfma_v8i e;
        e[0] = (int)((x.x >> sh) & M);
        e[1] = (int)((x.y >> sh) & M);
        e[2] = (int)((x.z >> sh) & M);
        e[3] = (int)((x.w >> sh) & M);
        e[4] = e[5] = e[6] = e[7] = 0;  // unused by FP4 operands
        return e;
      };
      mfma_v8i ea[2], en[2], eb[4];
#pragma unroll
      for (int fa = 0; fa < 2; ++fa) ea[fa] = ext(xa[fa], 0);
#pragma unroll
      for (int fb = 0; fb < 4; ++fb) eb[fb] = ext(xb[fb], 0);
      static_for<4>([&](auto PL) {
        constexpr int pl = decltype(PL)::value;
        constexpr int scale = pl < 3 ? kUnitScale + 2 - 2 * pl : kUnitScale;
        static_for<4>([&](auto FB) {
          constexpr int fb = decltype(FB)::value;
#pragma unroll
          for (int fa = 0; fa < 2; ++fa)
            acc[fa][fb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(ea[fa], eb[fb], acc[fa][fb], kFp4, kFp4,
                                                                          0, scale, 0, kUnitScale);
          if constexpr (pl < 3) {
            eb[fb] = ext(xb[fb], pl + 1);
            if constexpr (fb >= 2) en[fb - 2] = ext(xa[fb - 2], pl + 1);
          }
        });
        if constexpr (pl < 3) {
          ea[0] = en[0];
          ea[1] = en[1];
        }
      });
      } else {
      // bit planes 0..2 in a rolled loop (one plane's operands live at a time)
#pragma unroll 1
      for (int s = 0; s < 3; ++s) mfma_step(xa, xb, 0x11111111u << s, 0, kUnitScale + 2 - 2 * s);
      mfma_step(xa, xb, 0x22222222u, 2, kUnitScale);
      }
    }
  }

  const bool diag = bi == bj;
  const int rlim = n - row0, clim = n - col0;  // valid rows / columns of the tile
  const bool partial = diag || rlim < kTile || clim < kTile;  // block-uniform
  if constexpr (CAND) {
    // ---- candidate form: each accumulator value's bound compare is a lane
    // mask of the wave; a value with any survivor queues them (entry
    // li | lj' << 7 | popcount << 14), one atomic per flush reserves their
    // place in this block's output segment.  The queue is the wave's own
    // rows of the chunk buffer: after the main loop only this wave reads them.
    constexpr int kQueue = 64 * kRow;
    static_assert(kQueue >= 1024 + 256, "a fragment adds at most 16 x 64 entries");
    uint32_t *queue = Ps + wv * 64 * kRow;
    int qcount = 0;  // wave-uniform
    auto flush = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(counter, (unsigned long long)qcount);
      const int64_t b0 = (int64_t)__shfl(base, 0, 64);
      for (int e = lane; e < qcount; e += 64) {
        const uint32_t q = queue[e];
        const int64_t pos = b0 + e;
        if (pos < cap)
          out_keys[pos] = ((uint64_t)(row0 + (int)(q & 127u)) << kshift) | (uint64_t)(col0 + (int)((q >> 7) & 127u)) |
                          ((uint64_t)((q >> 14) & 0x1FFFu) << kCandPopShift);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      qcount = 0;
    };
    // PART: the tile is on the diagonal or the edge (i < j, i < n, j < n
    // checked per pair); otherwise the mask is the compare alone
    auto fragments = [&](auto PART) {
      constexpr bool P = decltype(PART)::value;
      static_for<2>([&](auto FA) {
        constexpr int fa = decltype(FA)::value;
        const int rb = wv * 64 + fa * 32 + 4 * g;  // rows rb + 8q + t hold accumulators 4q + t
        lds_f4 gr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) gr[q] = *reinterpret_cast<const lds_f4 *>(&sG[rb + 8 * q]);
        // the smallest row bound of each of the lane's 4 row groups
        // (rows rb + 8q .. rb + 8q + 3) and of all 16 (non-negative floats
        // order as their bits)
        uint32_t gq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          gq[q] = min(min(__float_as_uint(gr[q][0]), __float_as_uint(gr[q][1])),
                      min(__float_as_uint(gr[q][2]), __float_as_uint(gr[q][3])));
        const uint32_t grmin = min(min(gq[0], gq[1]), min(gq[2], gq[3]));
        static_for<4>([&](auto FB) {
          constexpr int fb = decltype(FB)::value;
          if (qcount > kQueue - 1024) flush();
          const int cj = fb * 32 + r32;  // the lane's column in the tile
          const float gc = sG[kTile + cj];
          const uint32_t gcu = __float_as_uint(gc);
          // a value can only be kept if its row group's largest popcount
          // reaches the group's smallest bound min(G_i, G_j): the whole
          // fragment is tested first (most fragments of a large run keep
          // nothing), then each row group (a fragment that keeps a pair
          // mostly keeps one), then the values of the groups that pass
          uint32_t amq[4];
#pragma unroll
          for (int q = 0; q < 4; ++q)
            amq[q] = max(max(__float_as_uint(acc[fa][fb][4 * q]), __float_as_uint(acc[fa][fb][4 * q + 1])),
                         max(__float_as_uint(acc[fa][fb][4 * q + 2]), __float_as_uint(acc[fa][fb][4 * q + 3])));
          const uint32_t amax = max(max(amq[0], amq[1]), max(amq[2], amq[3]));
          if (__ballot(amax >= min(grmin, gcu)) == 0) return;  // wave-uniform
          const uint32_t ebase = (uint32_t)rb | ((uint32_t)cj << 7);
          static_for<4>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            if (__ballot(amq[q] >= min(gq[q], gcu)) == 0) return;  // wave-uniform
            static_for<4>([&](auto T) {
              constexpr int r = 4 * q + decltype(T)::value;
              constexpr int dr = 8 * q + (r & 3);
              // popcount >= min(G_i, G_j): the pair may pass
              bool keep = acc[fa][fb][r] >= umin_f(gr[q][r & 3], gc);
              if constexpr (P) {
                const int ri = rb + dr;
                keep = keep && (!diag || ri < cj) && ri < rlim && cj < clim;
              }
              const uint64_t ball = __ballot(keep);
              if (ball) {  // wave-uniform
                if (keep) {  // slot: the kept lanes below this one
                  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0u));
                  queue[qcount + (int)below] = (ebase + (uint32_t)dr) | ((uint32_t)acc[fa][fb][r] << 14);
                }
                qcount += __popcll(ball);
              }
            });
          });
        });
      });
    };
    if (partial)
      fragments(std::true_type{});
    else
      fragments(std::false_type{});
    if (qcount > 0) flush();
    return;
  }
  // ---- bound: a 16-bit mask of possible pairs per fragment, branch-free
  uint32_t keep[8];
  int nkeep = 0;
  static_for<2>([&](auto FA) {
    constexpr int fa = decltype(FA)::value;
    const int rb = wv * 64 + fa * 32 + 4 * g;  // rows rb + 8q + t hold accumulators 4q + t
    lds_f4 gr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) gr[q] = *reinterpret_cast<const lds_f4 *>(&sG[rb + 8 * q]);
    static_for<4>([&](auto FB) {
      constexpr int fb = decltype(FB)::value;
      const int lj = kTile + fb * 32 + r32;
      const float gc = sG[lj];
      uint32_t rej = 0;
      static_for<16>([&](auto R) {
        constexpr int r = 15 - decltype(R)::value;  // bit r of the mask ends up holding pair r
        const float d = acc[fa][fb][r] - umin_f(gr[r >> 2][r & 3], gc);
        rej = __builtin_amdgcn_alignbit(rej, __float_as_uint(d), 31);  // rej << 1 | sign(d)
      });
      uint32_t k = ~rej & 0xFFFFu;
      if (partial) {  // i < j (row0 == col0 on the diagonal), i < n, j < n
        const int cj = lj - kTile;
        uint32_t ok = 0;
        static_for<16>([&](auto R) {
          constexpr int r = decltype(R)::value;
          const int ri = rb + 8 * (r >> 2) + (r & 3);
          ok |= (uint32_t)((!diag || ri < cj) && ri < rlim && cj < clim) << r;
        });
        k &= ok;
      }
      keep[fa * 4 + fb] = k;
      nkeep += __popc(k);
    });
  });

  // ---- block survivors; stage the repeat lists when they pay
  int wsum = nkeep;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off, 64);
  if (lane == 0) sWave[wv] = wsum;
  __syncthreads();  // also: both waves are done with the chunk buffer
  const int btot = sWave[0] + sWave[1];
  if (btot == 0) return;  // block-uniform
  bool staged = false;
  if (!CAND) {  // block-uniform
    const int la = min(kTile, n - row0) - 1, lb = min(kTile, n - col0) - 1;
    const int64_t a0 = sRo[0], b0 = sRo[kTile];
    const int lenA = (int)(sRo[la] + sRc[la] - a0);
    const int lenB = diag ? 0 : (int)(sRo[kTile + lb] + sRc[kTile + lb] - b0);
    staged = lenA + lenB <= kStageWords;
    if (staged) {
      for (int w = tid; w < lenA; w += 128) Ps[w] = rlist[a0 + w];
      for (int w = tid; w < lenB; w += 128) Ps[lenA + w] = rlist[b0 + w];
      for (int q = tid; q < 2 * kTile; q += 128)
        sLo[q] = q < kTile || diag ? (int)(sRo[q] - a0) : lenA + (int)(sRo[q] - b0);
    }
    __syncthreads();
  }

  // ---- survivors: queued per wave as li | lj' << 7 | popcount << 14 across
  // the fragments, then flushed: the lanes take the entries 64 apart, write the
  // exact count and the verdict back in place (li | lj' << 7 | m << 14 |
  // pass << 31; m <= 2048), one atomic per flush reserves the passing pairs and
  // the same lanes write them out.
  uint32_t *queue = sQ[wv];
  int qcount = 0;  // wave-uniform
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    int mine = 0;
    for (int e = lane; e < qcount; e += 64) {
      if (CAND) {
        ++mine;
      } else {
        const uint32_t q = queue[e];
        const int li = (int)(q & 127u), lj = kTile + (int)((q >> 7) & 127u);
        uint32_t m = q >> 14;
        const int na = sRc[li], nbr = sRc[lj];
        if (na > 0 && nbr > 0)
          m += staged ? repeat_min_sum(Ps + sLo[li], na, Ps + sLo[lj], nbr)
                      : repeat_min_sum(rlist + sRo[li], na, rlist + sRo[lj], nbr);
        const bool pass = passes(1.0f - (float)m / (float)min(sT[li], sT[lj]), thr);
        queue[e] = (q & 0x3FFFu) | (m << 14) | (pass ? 0x80000000u : 0u);
        mine += pass;
      }
    }
    const int incl = wave_incl_sum(mine, lane);
    const int total = __shfl(incl, 63, 64);
    if (total > 0) {  // wave-uniform
      unsigned long long base = 0;
      if (lane == 63) base = atomicAdd(counter, (unsigned long long)total);
      int64_t pos = (int64_t)__shfl(base, 63, 64) + (incl - mine);
      for (int e = lane; e < qcount; e += 64) {
        const uint32_t q = queue[e];
        if (CAND || (q >> 31)) {
          if (pos < cap) {
            const int li = (int)(q & 127u), lj = kTile + (int)((q >> 7) & 127u);
            out_keys[pos] = ((uint64_t)(row0 + li) << kshift) | (uint64_t)(col0 + lj - kTile) |
                            (CAND ? (uint64_t)((q >> 14) & 0x1FFFu) << kCandPopShift : 0ull);
            if (!CAND)
              out_dists[pos] = 1.0f - (float)((q >> 14) & 0x1FFFu) / (float)min(sT[li], sT[lj]);
          }
          ++pos;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    qcount = 0;
  };
  static_for<8>([&](auto F) {
    constexpr int fa = decltype(F)::value / 4, fb = decltype(F)::value % 4;
    const uint32_t k = keep[decltype(F)::value];
    if (__ballot(k != 0) == 0) return;  // wave-uniform
    const int c = __popc(k);
    const int incl = wave_incl_sum(c, lane);
    const int qn = __shfl(incl, 63, 64);
    if (qcount + qn > 1024) flush();
    int o = qcount + incl - c;
    static_for<16>([&](auto R) {
      constexpr int r = decltype(R)::value;
      if (k & (1u << r)) {
        const int li = wv * 64 + fa * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
        queue[o++] = (uint32_t)li | ((uint32_t)(fb * 32 + r32) << 7) | ((uint32_t)acc[fa][fb][r] << 14);
      }
    });
    qcount += qn;
  });
  if (qcount > 0) flush();
}

__global__ void unpack_keys_kernel(const uint64_t *__restrict__ keys, int64_t n, int kshift,
                                   int2 *__restrict__ pairs) {
  const uint64_t jmask = (1ull << kshift) - 1;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t v = keys[k];
    pairs[k] = make_int2((int)(v >> kshift), (int)(v & jmask));
  }
}

// ------------------------------------------------------ K = 7, 8 (exact)
// Every shared K-mer occurrence (K > 6) starts with a shared 6-mer at the same
// position, so matches_K <= I_K <= I_6 <= popcount_6 + min(excess_6): the
// k = 6 bitmap tiles, bounded against T_K = T_6 - (K - 6) and the K
// threshold, give a candidate superset (kmer_tile_mfma_kernel<6, true>).  The
// candidates then run the reference's merge (kmerCPU.cpp:29-43) over their
// K-mer lists sorted as signed short (kmerMain.cpp:156-159) -- its K = 8
// ordering included -- for the exact count.

// Ascending bitonic sort of the 64 R keys of one wave, R per lane: key
// e = R lane + q sits in v[q] of lane `lane`.  Stages with partner distance
// j >= R pair register q of lanes lane and lane ^ (j / R); shorter ones pair
// registers q and q + j of one lane.
template <int R>
__device__ __forceinline__ void bitonic_wave(uint32_t (&v)[R], const int lane) {
  constexpr int kLogP = (R == 8 ? 3 : 4) + 6;
#pragma unroll
  for (int lk = 1; lk <= kLogP; ++lk) {
    const int k = 1 << lk;
#pragma unroll
    for (int lj = lk - 1; lj >= 0; --lj) {
      const int j = 1 << lj;
      if (j >= R) {
        const int lx = j / R;
        const bool lower = (lane & lx) == 0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const uint32_t a = v[q];
          const uint32_t b = (uint32_t)__shfl_xor((int)a, lx, 64);
          const bool up = ((R * lane + q) & k) == 0;
          const bool keep_min = lower == up;
          v[q] = keep_min ? (a < b ? a : b) : (a < b ? b : a);
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (q & j) continue;
          const uint32_t a = v[q], b = v[q + j];
          const bool up = ((R * lane + q) & k) == 0;
          const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
          v[q] = up ? lo : hi;
          v[q + j] = up ? hi : lo;
        }
      }
    }
  }
}

// The T <= 64 R codes of one read sorted in one wave's registers: lane `lane`
// builds codes R lane .. R lane + R - 1 from two packed words (the R + K - 1
// bases they span start at base 8 lane or 16 lane, inside a 64-bit window).
template <int K, int R>
__device__ __forceinline__ void list_wave(const uint32_t *__restrict__ s, const int nw, const int T,
                                          const int lane, uint16_t *__restrict__ dst) {
  static_assert(R + K - 1 <= 24, "the lane's bases fit the 64-bit window past its offset");
  const int p0 = R * lane;
  const int w = p0 >> 4;
  const uint32_t lo = w < nw ? s[w] : 0u;
  const uint32_t hi = w + 1 < nw ? s[w + 1] : 0u;
  const uint64_t win = ((uint64_t)hi << 32 | lo) >> (2 * (p0 & 15));
  constexpr uint32_t kMask = (1u << (2 * K)) - 1u;
  uint32_t c = 0;
#pragma unroll
  for (int m = 0; m < K; ++m) c = (c << 2) | (uint32_t)((win >> (2 * m)) & 3u);
  uint32_t v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (q > 0) c = ((c << 2) | (uint32_t)((win >> (2 * (q + K - 1))) & 3u)) & kMask;
    // signed-short order of the code = unsigned order of code ^ 0x8000;
    // the padding sorts after every code
    v[q] = p0 + q < T ? (c ^ 0x8000u) : 0xFFFFFFFFu;
  }
  bitonic_wave<R>(v, lane);
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (p0 + q < T) dst[p0 + q] = (uint16_t)(v[q] ^ 0x8000u);
}

// One wave per read: the T = L-K+1 codes in signed-short order.  Up to 1,024
// codes sort in registers; longer reads sort in LDS (dynamic shared memory of
// CRISPY_MAX_READ_LEN keys, given only when some read needs it).
template <int K>
__global__ __launch_bounds__(64) void kmer_list_kernel(const uint32_t *__restrict__ pack,
                                                       const int64_t *__restrict__ poff,
                                                       const int32_t *__restrict__ len,
                                                       const int64_t *__restrict__ loff,
                                                       uint16_t *__restrict__ lists) {
  extern __shared__ uint32_t keys[];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const int L = len[r];
  const int T = L - K + 1;
  const uint32_t *s = pack + poff[r];
  uint16_t *dst = lists + loff[r];
  if (T <= 512) {
    list_wave<K, 8>(s, (L + 15) >> 4, T, tid, dst);
    return;
  }
  if (T <= 1024) {
    list_wave<K, 16>(s, (L + 15) >> 4, T, tid, dst);
    return;
  }
  int P = 1;
  while (P < T) P <<= 1;
  for (int t = tid; t < P; t += 64) {
    uint32_t c = 0xFFFFFFFFu;
    if (t < T) {
      c = 0;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const int p = t + q;
        c = (c << 2) | ((s[p >> 4] >> (2 * (p & 15))) & 3u);
      }
      c ^= 0x8000u;
    }
    keys[t] = c;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < P; t += 64) {
        const int o = t ^ j;
        if (o > t) {
          const uint32_t a = keys[t], b = keys[o];
          const bool up = (t & k) == 0;
          if ((a > b) == up) {
            keys[t] = b;
            keys[o] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = tid; t < T; t += 64) dst[t] = (uint16_t)(keys[t] ^ 0x8000u);
}

__global__ __launch_bounds__(256) void kmer_cand_kernel(
    const uint64_t *__restrict__ cand, int64_t ncand, const uint16_t *__restrict__ lists,
    const int64_t *__restrict__ loff, int K, const int32_t *__restrict__ len, float thr, int kshift,
    unsigned long long *__restrict__ counter, uint64_t *__restrict__ out_keys,
    float *__restrict__ out_dists, int64_t cap) {
  // Each lane merges its candidate's two lists through a window of kCandWin
  // codes of each, staged in LDS; the lanes of a wave refill their windows in
  // the same round, so the wave waits for memory once per round (each
  // round moves one list kCandWin - 3 codes or more) instead of whenever one
  // of its lanes crosses a word.  The windows
  // are stored lane-minor (code [list][position][thread]), so a merge step
  // reads its two codes with no unpacking and the lanes of a wave meet in a
  // bank at most in pairs whatever positions they are at.
  __shared__ uint16_t win[2][kCandWin][256];
  const int t = threadIdx.x, lane = t & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of outer iterations (shuffles)
  const int64_t base0 = (int64_t)blockIdx.x * blockDim.x + (t & ~63);
  for (int64_t b = base0; b < ncand; b += stride) {
    const int64_t c = b + lane;
    bool pass = false;
    float dist = 0.f;
    uint64_t key = 0;
    const uint16_t *La = lists, *Lb = lists;
    unsigned short l1 = 0, l2 = 0;
    if (c < ncand) {
      key = cand[c] & kCandKeyMask;
      const int i = (int)(key >> kshift), j = (int)(key & ((1ull << kshift) - 1));
      La = lists + loff[i];
      Lb = lists + loff[j];
      l1 = (unsigned short)(len[i] - K + 1);
      l2 = (unsigned short)(len[j] - K + 1);
    }
    unsigned short matches = 0;
    int k = 0, l = 0;
    bool active = c < ncand && l1 > 0 && l2 > 0;
    while (__ballot(active)) {
      if (active) {
        // the lists start at multiples of 4 codes: windows start on a word
        // (4 codes) and may read up to kListSlack codes past the last list
        const int ka = k & ~3, lb = l & ~3;
        const uint64_t *A = reinterpret_cast<const uint64_t *>(La + ka);
        const uint64_t *B = reinterpret_cast<const uint64_t *>(Lb + lb);
        uint64_t wa[kCandWin / 4], wb[kCandWin / 4];
#pragma unroll
        for (int q = 0; q < kCandWin / 4; ++q) {
          wa[q] = A[q];
          wb[q] = B[q];
        }
#pragma unroll
        for (int q = 0; q < kCandWin / 4; ++q)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            win[0][4 * q + s][t] = (uint16_t)(wa[q] >> (16 * s));
            win[1][4 * q + s][t] = (uint16_t)(wb[q] >> (16 * s));
          }
        const uint16_t *wA = &win[0][0][t], *wB = &win[1][0][t];
        int u = k - ka, v = l - lb;  // positions in the windows
        const int ue = min((int)l1, ka + kCandWin) - ka, ve = min((int)l2, lb + kCandWin) - lb;
        while (u < ue && v < ve) {
          const uint32_t x = wA[u * 256], y = wB[v * 256];
          matches = (unsigned short)(matches + (x == y));
          u += x <= y;
          v += x >= y;
        }
        k = ka + u;
        l = lb + v;
        active = k < l1 && l < l2;
      }
    }
    if (c < ncand) {
      const unsigned short mn = l1 < l2 ? l1 : l2;
      dist = 1.0f - (float)matches / (float)mn;
      pass = passes(dist, thr);
    }
    const uint64_t ball = __ballot(pass);
    const int total = __popcll(ball);
    unsigned long long basep = 0;
    if (total > 0) {
      if (lane == 0) basep = atomicAdd(counter, (unsigned long long)total);
      basep = __shfl(basep, 0, 64);
    }
    if (pass) {
      const int64_t pos = (int64_t)basep + __popcll(ball & ((1ull << lane) - 1));
      if (pos < cap) {
        out_keys[pos] = key;
        out_dists[pos] = dist;
      }
    }
  }
}

// Refines the candidates that the hashed k = 6 filter let through (their
// keys carry the bucket popcount): LPC lanes per candidate (32 or 16), lane l
// of a group ANDs words [128/LPC * l, 128/LPC * (l + 1)) of the two reads'
// 4,096-bit presence bitmaps for the exact presence popcount p, and the
// tile's exact-bitmap bound p + min(excess_i, excess_j) >=
// ((T_min - tshift) * cut - 1) keeps the candidates that might pass; those
// are queued per wave in LDS and appended to `out` with p in the popcount
// field (one atomic on *nout per up to 256 of them: a same-address atomic per
// wave step serialises across the chip).  A wave has 4 candidates' loads in
// flight per step (two per half wave, or one per quarter wave), with one
// memory round trip per step: the next keys are loaded a step ahead, and T
// and excess are requested with the bitmaps.  Half waves read each row as
// one 512-byte run per instruction (faster where candidates are sparse:
// reads250k 1.42 against 1.54 ms), quarter waves have less to reduce (faster
// where they are dense: reads100k 0.71 against 0.90 ms).
// Input: the tile pass's kCandSegs segments (seg_cap keys each, counts in
// `seg`); block b reads segment b mod kCandSegs, so the grid is a multiple of
// kCandSegs.
constexpr int kRefineQueue = 256;
template <int LPC>
__global__ __launch_bounds__(256) void kmer_cand_refine_kernel(
    const uint64_t *__restrict__ cand, const unsigned long long *__restrict__ seg, int64_t seg_cap,
    const uint32_t *__restrict__ bitmap, const int32_t *__restrict__ Tn, const int32_t *__restrict__ excess,
    float cut, int tshift, int kshift, unsigned long long *__restrict__ nout, uint64_t *__restrict__ out) {
  constexpr int WP = kmer_words(6);
  static_assert(LPC == 32 || LPC == 16, "half or quarter waves");
  constexpr int GW = 64 / LPC;   // candidate groups per wave
  constexpr int U = 4 / GW;      // candidates per group and step
  constexpr int WPL = WP / LPC;  // bitmap words per lane
  constexpr int NV = WPL / 4;    // 16-byte loads per lane and row
  static_assert(WPL % 4 == 0, "whole 16-byte loads");
  __shared__ uint64_t sq[4][kRefineQueue];
  const int lane = threadIdx.x & 63, h = lane / LPC, l = lane % LPC, wv = threadIdx.x >> 6;
  uint64_t *q = sq[wv];
  int qn = 0;  // wave-uniform
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(nout, (unsigned long long)qn);
    base = __shfl(base, 0, 64);
    for (int e = lane; e < qn; e += 64) out[base + e] = q[e];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    qn = 0;
  };
  const int sg = (int)(blockIdx.x % kCandSegs);
  const int64_t ncand = min((int64_t)seg[(size_t)sg * kSegStride], seg_cap);
  cand += (size_t)sg * seg_cap;
  const int64_t waves = (int64_t)(gridDim.x / kCandSegs) * (blockDim.x >> 6);
  const int64_t w0 = (int64_t)(blockIdx.x / kCandSegs) * (blockDim.x >> 6) + wv;
  uint64_t kn[U];
#pragma unroll
  for (int u = 0; u < U; ++u) kn[u] = 4 * w0 + GW * u + h < ncand ? cand[4 * w0 + GW * u + h] : 0;
  for (int64_t cb = 4 * w0; cb < ncand; cb += 4 * waves) {  // wave-uniform
    uint64_t key[U];
    bool valid[U];
    int p[U], Ti[U], Tj[U], Ei[U], Ej[U];
    uint4 a[U][NV], b[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {  // all next keys requested before any bitmap row
      const int64_t c = cb + GW * u + h;
      valid[u] = c < ncand;
      key[u] = valid[u] ? kn[u] & kCandKeyMask : 0;
      kn[u] = c + 4 * waves < ncand ? cand[c + 4 * waves] : 0;
      Ti[u] = Tj[u] = Ei[u] = Ej[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = (uint32_t)(key[u] >> kshift), j = (uint32_t)(key[u] & ((1ull << kshift) - 1));
#pragma unroll
      for (int v = 0; v < NV; ++v) a[u][v] = b[u][v] = make_uint4(0u, 0u, 0u, 0u);
      if (valid[u]) {
        const uint4 *ra = reinterpret_cast<const uint4 *>(bitmap + (size_t)i * WP + WPL * l);
        const uint4 *rb = reinterpret_cast<const uint4 *>(bitmap + (size_t)j * WP + WPL * l);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          a[u][v] = ra[v];
          b[u][v] = rb[v];
        }
        if (l == 0) {
          Ti[u] = Tn[i];
          Tj[u] = Tn[j];
          Ei[u] = excess[i];
          Ej[u] = excess[j];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      p[u] = 0;
#pragma unroll
      for (int v = 0; v < NV; ++v)
        p[u] += __popc(a[u][v].x & b[u][v].x) + __popc(a[u][v].y & b[u][v].y) + __popc(a[u][v].z & b[u][v].z) +
                __popc(a[u][v].w & b[u][v].w);
#pragma unroll
      for (int off = LPC / 2; off > 0; off >>= 1) p[u] += __shfl_xor(p[u], off, LPC);
      bool keep = false;
      if (valid[u] && l == 0) {
        const float F = fmaf((float)(min(Ti[u], Tj[u]) - tshift), cut, -1.0f);
        keep = (float)(p[u] + min(Ei[u], Ej[u])) >= F;
      }
      const uint64_t ball = __ballot(keep);
      if (keep) q[qn + __popcll(ball & ((1ull << lane) - 1))] = key[u] | ((uint64_t)p[u] << kCandPopShift);
      qn += __popcll(ball);
    }
    if (qn > kRefineQueue - 4) flush();
  }
  if (qn > 0) flush();
}

// Exact count of the K <= 6 candidates that the tile pass wrote out (pair key
// plus its presence popcount in the high bits): matches = popcount + the
// repeat-list merge (the same sum as in the tile's own survivor pass), the
// reference distance and predicate, the passing pairs appended.  Input:
// `nseg` segments of seg_cap keys whose counts are ncnt[s * kSegStride] (the
// tile pass's segments, or one run with its count at ncnt[0]); block b reads
// segment b mod nseg, so the grid is a multiple of nseg.  The repeat lists
// are merged a window of 8 entries per round trip, and the passing pairs are
// queued per wave in LDS and appended with one atomic on *counter per up to
// kExactQueue of them.
constexpr int kExactQueue = 256;
__global__ __launch_bounds__(256) void kmer_exact_kernel(
    const uint64_t *__restrict__ cand, const unsigned long long *__restrict__ ncnt, int nseg, int64_t seg_cap,
    const int32_t *__restrict__ Tn,
    const int32_t *__restrict__ rcnt, const int64_t *__restrict__ roffc, const uint32_t *__restrict__ rlistc,
    float thr, int kshift, unsigned long long *__restrict__ counter, uint64_t *__restrict__ out_keys,
    float *__restrict__ out_dists, int64_t cap) {
  __shared__ uint64_t sk[4][kExactQueue];
  __shared__ float sd[4][kExactQueue];
  const int sg = (int)(blockIdx.x % nseg);
  const int64_t ncand = min((int64_t)ncnt[(size_t)sg * kSegStride], seg_cap);
  cand += (size_t)sg * seg_cap;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int qn = 0;  // wave-uniform
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    unsigned long long basep = 0;
    if (lane == 0) basep = atomicAdd(counter, (unsigned long long)qn);
    basep = __shfl(basep, 0, 64);
    for (int e = lane; e < qn; e += 64) {
      const int64_t pos = (int64_t)basep + e;
      if (pos < cap) {
        out_keys[pos] = sk[wv][e];
        out_dists[pos] = sd[wv][e];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    qn = 0;
  };
  const int64_t stride = (int64_t)(gridDim.x / nseg) * blockDim.x;
  const int64_t base0 = (int64_t)(blockIdx.x / nseg) * blockDim.x + (threadIdx.x & ~63);
  for (int64_t b = base0; b < ncand; b += stride) {  // wave-uniform
    const int64_t c = b + lane;
    bool pass = false;
    float dist = 0.f;
    uint64_t key = 0;
    if (c < ncand) {
      const uint64_t e = cand[c];
      key = e & kCandKeyMask;
      const int i = (int)(key >> kshift), j = (int)(key & ((1ull << kshift) - 1));
      uint32_t m = (uint32_t)(e >> kCandPopShift);
      const int na = rcnt[i], nb = rcnt[j];
      if (na > 0 && nb > 0) m += repeat_min_sum_windowed<8>(rlistc + roffc[i], na, rlistc + roffc[j], nb);
      dist = 1.0f - (float)m / (float)min(Tn[i], Tn[j]);
      pass = passes(dist, thr);
    }
    const uint64_t ball = __ballot(pass);
    if (pass) {
      const int q = qn + __popcll(ball & ((1ull << lane) - 1));
      sk[wv][q] = key;
      sd[wv][q] = dist;
    }
    qn += __popcll(ball);
    if (qn > kExactQueue - 64) flush();
  }
  if (qn > 0) flush();
}

// The repeat lists back to back in read order: roffc = exclusive sum of the
// counts, one wave per read copies its list.
__global__ __launch_bounds__(256) void compact_repeats_kernel(const int32_t *__restrict__ rcnt,
                                                              const int64_t *__restrict__ roff,
                                                              const uint32_t *__restrict__ rlist,
                                                              const int64_t *__restrict__ roffc,
                                                              uint32_t *__restrict__ rlistc, int32_t n) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * 4;
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < n; r += waves) {
    const int c = rcnt[r];
    const uint32_t *src = rlist + roff[r];
    uint32_t *dst = rlistc + roffc[r];
    for (int e = lane; e < c; e += 64) dst[e] = src[e];
  }
}

// ------------------------------------------------------------ host driver
template <int K>
static int launch_profile(crispy_ctx *ctx, ReadStore &rs) {
  KmerState &km = ctx->km;
  if constexpr (K == 6) {
    if (km.hashed) {
      auto kernel = km.hash_words == 40   ? kmer_profile_kernel<6, 40>
                    : km.hash_words == 48 ? kmer_profile_kernel<6, 48>
                                          : kmer_profile_kernel<6, 64>;
      kernel<<<rs.n, 128, 0, ctx->stream>>>(
          rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(), km.bitmap.as<uint32_t>(),
          km.T.as<int32_t>(), km.rcnt.as<int32_t>(), km.roff.as<int64_t>(), km.rlist.as<uint32_t>(),
          km.excess.as<int32_t>(), km.hbitmap.as<uint32_t>(), km.hexcess.as<int32_t>());
      CRISPY_LAUNCHED(ctx);
      return CRISPY_OK;
    }
  }
  kmer_profile_kernel<K><<<rs.n, 128, 0, ctx->stream>>>(
      rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(), km.bitmap.as<uint32_t>(),
      km.T.as<int32_t>(), km.rcnt.as<int32_t>(), km.roff.as<int64_t>(),
      km.rlist.as<uint32_t>(), km.excess.as<int32_t>(), nullptr, nullptr);
  CRISPY_LAUNCHED(ctx);
  return CRISPY_OK;
}

// Zeroes a result counter (a kernel, not a memset: it is captured with the
// tile pass).
__global__ void zero_counter_kernel(unsigned long long *c) { *c = 0ull; }

// Zeroes the candidate segments' counters (one block of kCandSegs threads).
__global__ __launch_bounds__(kCandSegs) void zero_segs_kernel(unsigned long long *seg) {
  seg[(size_t)threadIdx.x * kSegStride] = 0ull;
}

// stat[0] = the segments' total count, stat[1] = the largest (one block of
// kCandSegs threads).
__global__ __launch_bounds__(kCandSegs) void seg_stat_kernel(const unsigned long long *__restrict__ seg,
                                                             unsigned long long *__restrict__ stat) {
  static_assert(kCandSegs % 64 == 0 && kCandSegs <= 1024, "whole waves, one block");
  __shared__ unsigned long long ssum[kCandSegs / 64], smax[kCandSegs / 64];
  unsigned long long c = seg[(size_t)threadIdx.x * kSegStride], m = c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off, 64);
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) {
    ssum[threadIdx.x >> 6] = c;
    smax[threadIdx.x >> 6] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, x = 0;
    for (int w = 0; w < kCandSegs / 64; ++w) {
      t += ssum[w];
      x = smax[w] > x ? smax[w] : x;
    }
    stat[0] = t;
    stat[1] = x;
  }
}

// Read pairs i < j < n inside tile slots [tb, tb + nt): the tile pass's units.
static double tile_units(int32_t n, int nb, int64_t tb, int64_t nt) {
  if (tb == 0 && nt == num_slots(nb)) return (double)n * (n - 1) / 2;
  double units = 0;
  for (int64_t t = tb; t < tb + nt; ++t) {
    int bi, bj;
    if (!slot_coords(t, nb, bi, bj)) continue;
    const int64_t rh = std::min<int64_t>(kTile, n - (int64_t)bi * kTile);
    const int64_t chh = std::min<int64_t>(kTile, n - (int64_t)bj * kTile);
    units += bi == bj ? (double)rh * (rh - 1) / 2 : (double)rh * chh;
  }
  return units;
}

// The tile pass: zero the result counter, then the tile kernel over slots
// [tb, tb + nt) (a captured graph with graphs on).  HW > 0: the candidate
// form on the k = 6 bucket bitmaps (HW words per read) and their excess;
// otherwise the presence bitmaps of K.  CAND: the output is kCandSegs
// segments of `cap` keys (counters in km.seg, their total and largest count
// in km.seg_stat after the pass).
template <int K, bool CAND, int HW = 0>
static int launch_tiles(crispy_ctx *ctx, int32_t n, int nb, int64_t tb, int64_t nt, float thr,
                        int tshift, uint64_t *out_keys, float *out_dists, int64_t cap) {
  constexpr bool HASHED = HW > 0;
  static_assert(!HASHED || (K == 6 && CAND), "the hashed filter is the k = 6 candidate form");
  constexpr int WP = HASHED ? HW : kmer_words(K);
  constexpr int CW = CAND && (WP == 40 || WP == 48) ? 24 : kChunkWords;  // chunk words (2 chunks per tile)
  KmerState &km = ctx->km;
  const DevBuf &bits = HASHED ? km.hbitmap : km.bitmap;
  const DevBuf &excess = HASHED ? km.hexcess : km.excess;
  const int64_t per_xcd = (nt + 7) / 8;
  const int64_t grid = per_xcd * 8;
  const float cut = 1.0f - thr - 2e-5f;
  uint32_t thr_bits;
  std::memcpy(&thr_bits, &thr, sizeof thr_bits);
  unsigned long long *counter = CAND ? km.seg.as<unsigned long long>() : km.counter.as<unsigned long long>();
  hipEvent_t ev;
  timer_begin(ctx, "kmer_tile", &ev);
  CRISPY_TRY(run_segment(
      ctx, km.g_tile,
      {(uintptr_t)WP, (uintptr_t)CAND, (uintptr_t)bits.p, (uintptr_t)km.T.p, (uintptr_t)km.rcnt.p,
       (uintptr_t)km.roffc.p, (uintptr_t)km.rlistc.p, (uintptr_t)excess.p, (uintptr_t)n, (uintptr_t)nb,
       (uintptr_t)tb, (uintptr_t)nt, thr_bits, (uintptr_t)tshift, (uintptr_t)km.kshift, (uintptr_t)counter,
       (uintptr_t)km.seg_stat.p, (uintptr_t)out_keys, (uintptr_t)out_dists, (uintptr_t)cap},
      "k-mer tile pass", [&](bool) -> int {
        if (CAND)
          zero_segs_kernel<<<1, kCandSegs, 0, ctx->stream>>>(counter);
        else
          zero_counter_kernel<<<1, 1, 0, ctx->stream>>>(counter);
        CRISPY_LAUNCHED(ctx);
        kmer_tile_mfma_kernel<WP, CAND, CW><<<(unsigned)grid, 128, 0, ctx->stream>>>(
            bits.as<uint32_t>(), km.T.as<int32_t>(), km.rcnt.as<int32_t>(), km.roffc.as<int64_t>(),
            km.rlistc.as<uint32_t>(), excess.as<int32_t>(), n, nb, tb, nt, per_xcd, thr, cut, tshift,
            km.kshift, counter, out_keys, out_dists, cap);
        CRISPY_LAUNCHED(ctx);
        if (CAND) {
          seg_stat_kernel<<<1, kCandSegs, 0, ctx->stream>>>(counter, km.seg_stat.as<unsigned long long>());
          CRISPY_LAUNCHED(ctx);
        }
        return CRISPY_OK;
      }));
  timer_end(ctx, "kmer_tile", ev, ctx->timing ? tile_units(n, nb, tb, nt) : 0.0);
  return CRISPY_OK;
}

using ProfileFn = int (*)(crispy_ctx *, ReadStore &);
using TileFn = int (*)(crispy_ctx *, int32_t, int, int64_t, int64_t, float, int, uint64_t *, float *,
                       int64_t);
static const ProfileFn kProfile[7] = {nullptr,           launch_profile<1>, launch_profile<2>,
                                      launch_profile<3>, launch_profile<4>, launch_profile<5>,
                                      launch_profile<6>};
static const TileFn kTiles[7] = {nullptr,
                                 launch_tiles<1, false>,
                                 launch_tiles<2, false>,
                                 launch_tiles<3, false>,
                                 launch_tiles<4, false>,
                                 launch_tiles<5, false>,
                                 launch_tiles<6, false>};
static const TileFn kTilesCand[7] = {nullptr,
                                     launch_tiles<1, true>,
                                     launch_tiles<2, true>,
                                     launch_tiles<3, true>,
                                     launch_tiles<4, true>,
                                     launch_tiles<5, true>,
                                     launch_tiles<6, true>};

// Profiles of the bitmap word size KP (KP = K for K <= 6, 6 above); every
// read must hold at least one K-mer.
static int build_profiles(crispy_ctx *ctx, int KP, int K, bool hash) {
  ReadStore &rs = ctx->reads[CRISPY_READS_KMER];
  KmerState &km = ctx->km;
  km.hashed = hash && KP == 6;
  km.hash_words = km.hash_buckets ? km.hash_buckets / 32 : kDefaultHashWords;
  // Rebuilt on every run: the per-read profiles are part of the stage's work.
  const int32_t n = rs.n;
  for (int32_t r = 0; r < n; ++r)
    if (rs.h_len[r] - K + 1 < 1)
      return set_err(ctx, CRISPY_ERR_ARG, "read %d is shorter than K=%d (length %d)", r, K, rs.h_len[r]);
  // list offsets (and, for K > 6, the K-mer list offsets) staged in pinned
  // memory, so their uploads need no wait
  const size_t offs = sizeof(int64_t) * (n + 1);
  CRISPY_TRY(ensure_host(ctx, K > 6 ? 2 * offs : offs));
  int64_t *roff = static_cast<int64_t *>(ctx->h_stage);
  roff[0] = 0;
  for (int32_t r = 0; r < n; ++r) roff[r + 1] = roff[r] + (rs.h_len[r] - KP + 1) / 2 + 1;
  const int WP = kmer_words(KP);
  CRISPY_TRY(ensure(ctx, km.bitmap, sizeof(uint32_t) * (size_t)n * WP));
  CRISPY_TRY(ensure(ctx, km.T, sizeof(int32_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, km.rcnt, sizeof(int32_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, km.excess, sizeof(int32_t) * (n + 1)));
  if (km.hashed) {
    // + kBitmapSlack: the tile's last part-filled chunk is staged whole
    CRISPY_TRY(ensure(ctx, km.hbitmap, sizeof(uint32_t) * ((size_t)n * km.hash_words + kBitmapSlack)));
    CRISPY_TRY(ensure(ctx, km.hexcess, sizeof(int32_t) * (n + 1)));
  }
  CRISPY_TRY(ensure(ctx, km.roff, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, km.rlist, sizeof(uint32_t) * (roff[n] + 1)));
  CRISPY_HIP(ctx, hipMemcpyAsync(km.roff.p, roff, offs, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_TRY(ensure(ctx, km.roffc, sizeof(int64_t) * (n + 1)));
  CRISPY_TRY(ensure(ctx, km.rlistc, sizeof(uint32_t) * (roff[n] + 1)));
  if (K > 6) {  // the K-mer lists for the exact merge of the candidates
    int64_t *loff = roff + (n + 1);
    loff[0] = 0;
    for (int32_t r = 0; r < n; ++r) loff[r + 1] = loff[r] + ((rs.h_len[r] - K + 1 + 3) & ~3);  // 8-byte aligned
    CRISPY_TRY(ensure(ctx, km.loff, offs));
    CRISPY_TRY(ensure(ctx, km.lists, sizeof(uint16_t) * (loff[n] + kListSlack)));
    CRISPY_HIP(ctx, hipMemcpyAsync(km.loff.p, loff, offs, hipMemcpyHostToDevice, ctx->stream));
  }
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::exclusive_scan(nullptr, tbytes, km.rcnt.as<int32_t>(), km.roffc.as<int64_t>(),
                                          (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  // The profile pass (a captured graph with graphs on): profiles, the
  // repeat lists compacted in read order, and for K > 6 the K-mer lists.
  // the K-mer list kernel sorts reads of more than 1,024 K-mers in LDS
  const size_t list_lds = rs.maxlen - K + 1 > 1024 ? sizeof(uint32_t) * CRISPY_MAX_READ_LEN : 0;
  hipEvent_t ev;
  timer_begin(ctx, "kmer_profile", &ev);
  CRISPY_TRY(run_segment(
      ctx, km.g_profile,
      {(uintptr_t)KP, (uintptr_t)K, (uintptr_t)n, (uintptr_t)rs.pack.p, (uintptr_t)rs.poff.p, (uintptr_t)rs.len.p,
       (uintptr_t)km.bitmap.p, (uintptr_t)km.T.p, (uintptr_t)km.rcnt.p, (uintptr_t)km.roff.p, (uintptr_t)km.rlist.p,
       (uintptr_t)km.excess.p, (uintptr_t)km.roffc.p, (uintptr_t)km.rlistc.p, (uintptr_t)ctx->temp.p,
       (uintptr_t)tbytes, (uintptr_t)(K > 6 ? km.loff.p : nullptr), (uintptr_t)(K > 6 ? km.lists.p : nullptr),
       (uintptr_t)list_lds, (uintptr_t)(km.hashed ? km.hash_words : 0),
       (uintptr_t)(km.hashed ? km.hbitmap.p : nullptr),
       (uintptr_t)(km.hashed ? km.hexcess.p : nullptr)},
      "k-mer profile pass", [&](bool) -> int {
        CRISPY_TRY(kProfile[KP](ctx, rs));
        size_t tb = tbytes;
        CRISPY_HIP(ctx, rocprim::exclusive_scan(ctx->temp.p, tb, km.rcnt.as<int32_t>(), km.roffc.as<int64_t>(),
                                                (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
        compact_repeats_kernel<<<(int)std::min<int64_t>((n + 3) / 4, 4096), 256, 0, ctx->stream>>>(
            km.rcnt.as<int32_t>(), km.roff.as<int64_t>(), km.rlist.as<uint32_t>(), km.roffc.as<int64_t>(),
            km.rlistc.as<uint32_t>(), n);
        CRISPY_LAUNCHED(ctx);
        if (K == 7)
          kmer_list_kernel<7><<<n, 64, list_lds, ctx->stream>>>(rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(),
                                                               rs.len.as<int32_t>(), km.loff.as<int64_t>(),
                                                               km.lists.as<uint16_t>());
        else if (K == 8)
          kmer_list_kernel<8><<<n, 64, list_lds, ctx->stream>>>(rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(),
                                                               rs.len.as<int32_t>(), km.loff.as<int64_t>(),
                                                               km.lists.as<uint16_t>());
        CRISPY_LAUNCHED(ctx);
        return CRISPY_OK;
      }));
  timer_end(ctx, "kmer_profile", ev, n);
  CRISPY_HIP(ctx, hipEventRecord(ctx->stage_ev, ctx->stream));
  km.K = K;
  km.profile_gen = rs.generation;
  return CRISPY_OK;
}

// Runs `launch` (which zeroes km.counter, then appends to it) and grows the
// output once if it overflowed `cap`; returns the final count.
template <class F>
static int run_counted(crispy_ctx *ctx, int64_t &cap, DevBuf &keys, DevBuf *dists, int64_t *count,
                       F &&launch) {
  KmerState &km = ctx->km;
  CRISPY_TRY(ensure(ctx, km.counter, sizeof(unsigned long long)));
  for (int attempt = 0; attempt < 2; ++attempt) {
    CRISPY_TRY(ensure(ctx, keys, sizeof(uint64_t) * cap));
    if (dists) CRISPY_TRY(ensure(ctx, *dists, sizeof(float) * cap));
    CRISPY_TRY(launch());
    unsigned long long c = 0;
    CRISPY_HIP(ctx, hipMemcpyAsync(&c, km.counter.p, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = (int64_t)c;
    if (*count <= cap) return CRISPY_OK;
    if (attempt == 0) cap = *count + *count / 8 + 1024;  // rerun with room for every entry
  }
  // the rerun produced more entries than its own first count: the buffers
  // hold only `cap` of them, so nothing downstream may read `*count`
  return set_err(ctx, CRISPY_ERR_STATE, "k-mer output overflowed twice (%lld entries, room for %lld)",
                 (long long)*count, (long long)cap);
}

// The candidate tile pass (`launch` zeroes the segment counters, appends to
// kCandSegs segments of km.cand_seg_cap keys and totals them in km.seg_stat):
// reruns once with larger segments if one overflowed; returns the total.
template <class F>
static int run_segmented(crispy_ctx *ctx, int64_t *count, F &&launch) {
  KmerState &km = ctx->km;
  CRISPY_TRY(ensure(ctx, km.seg, sizeof(unsigned long long) * kCandSegs * kSegStride));
  CRISPY_TRY(ensure(ctx, km.seg_stat, 2 * sizeof(unsigned long long)));
  unsigned long long st[2] = {0, 0};
  for (int attempt = 0; attempt < 2; ++attempt) {
    CRISPY_TRY(ensure(ctx, km.cand, sizeof(uint64_t) * kCandSegs * km.cand_seg_cap));
    CRISPY_TRY(launch());
    CRISPY_HIP(ctx, hipMemcpyAsync(st, km.seg_stat.p, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = (int64_t)st[0];
    if ((int64_t)st[1] <= km.cand_seg_cap) return CRISPY_OK;
    if (attempt == 0) km.cand_seg_cap = (int64_t)st[1] + (int64_t)st[1] / 8 + 256;  // room for every segment
  }
  return set_err(ctx, CRISPY_ERR_STATE, "k-mer candidates overflowed a segment twice (%llu, room for %lld)",
                 st[1], (long long)km.cand_seg_cap);
}

// The hashed filter's candidates (the tile pass's segments) refined by the
// exact k = 6 bitmaps into km.cand_alt, their count left in km.cand_n.
static int refine_candidates(crispy_ctx *ctx, int64_t ncand, double range_pairs, float threshold, int tshift) {
  KmerState &km = ctx->km;
  CRISPY_TRY(ensure(ctx, km.cand_alt, sizeof(uint64_t) * std::max<int64_t>(ncand, 1)));
  CRISPY_TRY(ensure(ctx, km.cand_n, sizeof(unsigned long long)));
  hipEvent_t ev;
  timer_begin(ctx, "kmer_cand_pop", &ev);
  zero_counter_kernel<<<1, 1, 0, ctx->stream>>>(km.cand_n.as<unsigned long long>());
  CRISPY_LAUNCHED(ctx);
  // 16 candidates per block step; at most 8 blocks per segment.  Quarter
  // waves from 1 candidate per 1,000 pairs of the range up (about 16 per tile)
  const int64_t per_seg = (ncand + kCandSegs - 1) / kCandSegs;
  const int bps = (int)std::max<int64_t>(1, std::min<int64_t>(8, (per_seg + 63) / 64));
  const auto kernel = (double)ncand >= 1e-3 * range_pairs ? kmer_cand_refine_kernel<16> : kmer_cand_refine_kernel<32>;
  kernel<<<kCandSegs * bps, 256, 0, ctx->stream>>>(
      km.cand.as<uint64_t>(), km.seg.as<unsigned long long>(), km.cand_seg_cap, km.bitmap.as<uint32_t>(),
      km.T.as<int32_t>(), km.excess.as<int32_t>(), 1.0f - threshold - 2e-5f, tshift, km.kshift,
      km.cand_n.as<unsigned long long>(), km.cand_alt.as<uint64_t>());
  CRISPY_LAUNCHED(ctx);
  timer_end(ctx, "kmer_cand_pop", ev, (double)ncand);
  return CRISPY_OK;
}

static int sort_keys(crispy_ctx *ctx, int64_t count) {
  KmerState &km = ctx->km;
  if (count <= 1) {
    if (count == 1) {
      CRISPY_HIP(ctx, hipMemcpyAsync(km.keys_alt.p, km.keys.p, 8, hipMemcpyDeviceToDevice, ctx->stream));
      CRISPY_HIP(ctx, hipMemcpyAsync(km.dists_alt.p, km.dists.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return CRISPY_OK;
  }
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(nullptr, tbytes, km.keys.as<uint64_t>(),
                                            km.keys_alt.as<uint64_t>(), km.dists.as<float>(),
                                            km.dists_alt.as<float>(), (size_t)count, 0, 2 * km.kshift,
                                            ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  hipEvent_t ev;
  timer_begin(ctx, "kmer_sort", &ev);
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs<RadixOnesweep>(ctx->temp.p, tbytes, km.keys.as<uint64_t>(),
                                            km.keys_alt.as<uint64_t>(), km.dists.as<float>(),
                                            km.dists_alt.as<float>(), (size_t)count, 0, 2 * km.kshift,
                                            ctx->stream));
  timer_end(ctx, "kmer_sort", ev, (double)count);
  return CRISPY_OK;
}

}  // namespace crispy

using namespace crispy;

extern "C" {

int64_t crispy_kmer_num_tiles(int32_t n) { return num_slots((n + kTile - 1) / kTile); }

int crispy_kmer_tile_bits(crispy_ctx *ctx) {
  if (!ctx || ctx->km.K == 0) return 0;
  return 32 * (ctx->km.hashed ? ctx->km.hash_words : kmer_words(ctx->km.K <= 6 ? ctx->km.K : 6));
}

int crispy_kmer_set_exact_pass(crispy_ctx *ctx, int mode) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (mode < -1 || mode > 2) return set_err(ctx, CRISPY_ERR_ARG, "exact pass mode %d: -1, 0, 1 or 2", mode);
  ctx->km.exact_pass = mode;
  return CRISPY_OK;
}

int crispy_kmer_set_hash_buckets(crispy_ctx *ctx, int buckets) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (buckets != 0 && (buckets % 32 != 0 || !valid_hash_words(buckets / 32)))
    return set_err(ctx, CRISPY_ERR_ARG, "hash buckets %d: 0 (default), 1280, 1536 or 2048", buckets);
  ctx->km.hash_buckets = buckets;
  return CRISPY_OK;
}

int crispy_kmer_autotune(crispy_ctx *ctx, int K, float threshold, int repeats, int *chosen) {
  if (!ctx || repeats < 1) return CRISPY_ERR_ARG;
  if (K < 1 || K > 8) return set_err(ctx, CRISPY_ERR_ARG, "unsupported K=%d (1..8)", K);
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  // K <= 5 have no hashed form (2 = 1); K > 6 always take candidates (0 = 1);
  // the hashed form (mode 1, K >= 6) with each bucket count
  std::vector<std::pair<int, int>> settings;  // (exact pass mode, hash buckets)
  if (K <= 6) settings.push_back({0, 0});
  if (K >= 6) {
    for (int buckets : {2048, 1536, 1280}) settings.push_back({1, buckets});
    settings.push_back({2, 0});
  } else {
    settings.push_back({1, 0});
  }
  KmerState &km = ctx->km;
  const std::pair<int, int> before{km.exact_pass, km.hash_buckets};
  double best = 1e300;
  std::pair<int, int> pick = before;
  for (const auto &st : settings) {
    km.exact_pass = st.first;
    km.hash_buckets = st.second;
    double fastest = 1e300;
    for (int r = 0; r <= repeats; ++r) {  // the first run warms up (profiles, graphs, buffers)
      CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = crispy_kmer_run(ctx, K, threshold, 0, -1, nullptr);
      if (rc != CRISPY_OK) {
        km.exact_pass = before.first;
        km.hash_buckets = before.second;
        return rc;
      }
      CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (r > 0) fastest = std::min(fastest, sec);
    }
    if (fastest < best) {
      best = fastest;
      pick = st;
    }
  }
  km.exact_pass = pick.first;
  km.hash_buckets = pick.second;
  // leave the result of the chosen setting (every setting gives the same one)
  CRISPY_TRY(crispy_kmer_run(ctx, K, threshold, 0, -1, nullptr));
  if (chosen) *chosen = pick.first;
  return CRISPY_OK;
}

int crispy_kmer_run(crispy_ctx *ctx, int K, float threshold, int64_t tile_begin, int64_t tile_end,
                    int64_t *count) {
  if (!ctx) return CRISPY_ERR_ARG;
  if (K < 1 || K > 8) return set_err(ctx, CRISPY_ERR_ARG, "unsupported K=%d (1..8)", K);
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  ReadStore &rs = ctx->reads[CRISPY_READS_KMER];
  KmerState &km = ctx->km;
  km.valid = false;
  ctx->nw.valid = false;
  const int32_t n = rs.n;
  if ((int64_t)n * kmer_words(6) > (int64_t)UINT32_MAX)  // the tiles address bitmaps with 32-bit word offsets
    return set_err(ctx, CRISPY_ERR_ARG, "%d reads: at most %lld per k-mer run", n,
                   (long long)(UINT32_MAX / kmer_words(6)));
  const int nb = (n + kTile - 1) / kTile;
  const int64_t total_tiles = num_slots(nb);
  if (tile_end < 0 || tile_end > total_tiles) tile_end = total_tiles;
  if (tile_begin < 0 || tile_begin > tile_end)
    return set_err(ctx, CRISPY_ERR_ARG, "bad tile range [%lld, %lld)", (long long)tile_begin,
                   (long long)tile_end);
  const int64_t nt = tile_end - tile_begin;
  km.count = 0;
  if (n < 2 || nt == 0) {
    km.valid = true;
    if (count) *count = 0;
    return CRISPY_OK;
  }
  // pair keys i << kshift | j with 2^kshift >= n: as few radix-sort bits as the read count needs
  km.kshift = 1;
  while ((1 << km.kshift) < n) ++km.kshift;
  const int KP = K <= 6 ? K : 6;
  // K <= 6: exact counts in the tile, or in a pass of their own over the
  // tile's candidates.  Measured on MI355X: the lean candidate tile plus the
  // pass wins from ~10^9 pairs up (reads100k 9.8 -> 9.1 ms, reads250k 50.2
  // -> 44.0 ms per k-mer run); below, the pass's latency-bound merges cost more
  // than the tile saves (reads20k 0.78 -> 0.97 ms).  The candidate tiles of
  // k = 6 run on the hashed bitmaps (half the words).
  // exact_pass 2: the candidate form on the exact bitmaps (no hashing).
  const bool separate = km.exact_pass >= 1 || (km.exact_pass == -1 && (int64_t)n * (n - 1) / 2 >= kExactPassPairs);
  CRISPY_TRY(build_profiles(ctx, KP, K, (K > 6 || separate) && km.exact_pass != 2));
  // the k = 6 candidate tiles: on the hashed bitmaps, or on the exact ones
  const TileFn tiles6 = !km.hashed                          ? launch_tiles<6, true>
                        : km.hash_words == 40            ? launch_tiles<6, true, 40>
                        : km.hash_words == 48            ? launch_tiles<6, true, 48>
                                                         : launch_tiles<6, true, 64>;
  // first guess at the result size: the pairs in the range, at most 4 Mi
  const int64_t pairs_in_range = std::min<int64_t>(nt * kTile * kTile, (int64_t)n * (n - 1) / 2);
  const int64_t guess = std::min<int64_t>((int64_t)1 << 22, std::max<int64_t>(4096, pairs_in_range));
  km.cap = std::max(km.cap, guess);
  if (K <= 6 && !separate) {
    CRISPY_TRY(run_counted(ctx, km.cap, km.keys, &km.dists, &km.count, [&]() {
      return kTiles[K](ctx, n, nb, tile_begin, nt, threshold, 0, km.keys.as<uint64_t>(),
                       km.dists.as<float>(), km.cap);
    }));
  } else if (K <= 6) {  // the tile pass writes candidates, a second pass counts them exactly
    km.cand_seg_cap = std::max(km.cand_seg_cap, (guess + kCandSegs - 1) / kCandSegs);
    int64_t ncand = 0;
    CRISPY_TRY(run_segmented(ctx, &ncand, [&]() {
      return (K == 6 ? tiles6 : kTilesCand[K])(ctx, n, nb, tile_begin, nt, threshold, 0,
                                               km.cand.as<uint64_t>(), nullptr, km.cand_seg_cap);
    }));
    km.candidates = ncand;
    // hashed filter (K = 6): the exact presence popcount and the exact-bitmap
    // bound first, so that only the candidates the tile would have kept with
    // the exact bitmaps go on to the repeat-list merge (one run, its count on
    // the device); otherwise the merge reads the tile's segments
    const uint64_t *exact_in = km.cand.as<uint64_t>();
    const unsigned long long *exact_n = km.seg.as<unsigned long long>();
    int nseg = kCandSegs;
    int64_t seg_cap = km.cand_seg_cap;
    if (km.hashed && ncand > 0) {
      CRISPY_TRY(refine_candidates(ctx, ncand, (double)pairs_in_range, threshold, 0));
      exact_in = km.cand_alt.as<uint64_t>();
      exact_n = km.cand_n.as<unsigned long long>();
      nseg = 1;
      seg_cap = ncand;
    }
    CRISPY_TRY(run_counted(ctx, km.cap, km.keys, &km.dists, &km.count, [&]() -> int {
      zero_counter_kernel<<<1, 1, 0, ctx->stream>>>(km.counter.as<unsigned long long>());
      CRISPY_LAUNCHED(ctx);
      if (ncand == 0) return CRISPY_OK;
      hipEvent_t ev;
      timer_begin(ctx, "kmer_cand", &ev);
      const int64_t per_seg = (ncand + nseg - 1) / nseg;
      const int bps = (int)std::max<int64_t>(1, std::min<int64_t>(8192 / nseg, (per_seg + 255) / 256));
      kmer_exact_kernel<<<nseg * bps, 256, 0, ctx->stream>>>(
          exact_in, exact_n, nseg, seg_cap, km.T.as<int32_t>(), km.rcnt.as<int32_t>(), km.roffc.as<int64_t>(),
          km.rlistc.as<uint32_t>(), threshold, km.kshift, km.counter.as<unsigned long long>(),
          km.keys.as<uint64_t>(), km.dists.as<float>(), km.cap);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "kmer_cand", ev, (double)ncand);
      return CRISPY_OK;
    }));
  } else {
    km.cand_seg_cap = std::max(km.cand_seg_cap, (guess + kCandSegs - 1) / kCandSegs);
    int64_t ncand = 0;
    CRISPY_TRY(run_segmented(ctx, &ncand, [&]() {
      return tiles6(ctx, n, nb, tile_begin, nt, threshold, K - 6, km.cand.as<uint64_t>(), nullptr,
                    km.cand_seg_cap);
    }));
    km.candidates = ncand;
    // the exact k = 6 bitmap bound against T_K = T_6 - (K - 6), into one run
    // (on exact-bitmap candidates it keeps them all: then it only compacts)
    int64_t nref = 0;
    if (ncand > 0) {
      CRISPY_TRY(refine_candidates(ctx, ncand, (double)pairs_in_range, threshold, K - 6));
      unsigned long long c = 0;
      CRISPY_HIP(ctx, hipMemcpyAsync(&c, km.cand_n.p, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
      CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
      nref = (int64_t)c;
    }
    // (i, j) order, so the lanes of a wave merge against the same few row lists
    const uint64_t *cand = km.cand_alt.as<uint64_t>();
    if (nref > 1) {
      size_t tbytes = 0;
      CRISPY_HIP(ctx, rocprim::radix_sort_keys<RadixOnesweep>(nullptr, tbytes, km.cand_alt.as<uint64_t>(),
                                               km.cand.as<uint64_t>(), (size_t)nref, 0, 2 * km.kshift,
                                               ctx->stream));
      CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
      hipEvent_t sev;
      timer_begin(ctx, "kmer_cand_sort", &sev);
      CRISPY_HIP(ctx, rocprim::radix_sort_keys<RadixOnesweep>(ctx->temp.p, tbytes, km.cand_alt.as<uint64_t>(),
                                               km.cand.as<uint64_t>(), (size_t)nref, 0, 2 * km.kshift,
                                               ctx->stream));
      timer_end(ctx, "kmer_cand_sort", sev, (double)nref);
      cand = km.cand.as<uint64_t>();
    }
    CRISPY_TRY(run_counted(ctx, km.cap, km.keys, &km.dists, &km.count, [&]() -> int {
      zero_counter_kernel<<<1, 1, 0, ctx->stream>>>(km.counter.as<unsigned long long>());
      CRISPY_LAUNCHED(ctx);
      if (nref == 0) return CRISPY_OK;
      hipEvent_t ev;
      timer_begin(ctx, "kmer_cand", &ev);
      const int blocks = (int)std::min<int64_t>((nref + 255) / 256, 8192);
      kmer_cand_kernel<<<blocks, 256, 0, ctx->stream>>>(
          cand, nref, km.lists.as<uint16_t>(), km.loff.as<int64_t>(), K,
          rs.len.as<int32_t>(), threshold, km.kshift, km.counter.as<unsigned long long>(), km.keys.as<uint64_t>(),
          km.dists.as<float>(), km.cap);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "kmer_cand", ev, (double)nref);
      return CRISPY_OK;
    }));
  }
  CRISPY_TRY(ensure(ctx, km.keys_alt, sizeof(uint64_t) * (km.count + 1)));
  CRISPY_TRY(ensure(ctx, km.dists_alt, sizeof(float) * (km.count + 1)));
  CRISPY_TRY(ensure(ctx, km.pairs, sizeof(int2) * (km.count + 1)));
  CRISPY_TRY(sort_keys(ctx, km.count));
  if (km.count > 0) {
    const int blocks = (int)std::min<int64_t>((km.count + 255) / 256, 4096);
    unpack_keys_kernel<<<blocks, 256, 0, ctx->stream>>>(km.keys_alt.as<uint64_t>(), km.count,
                                                        km.kshift, km.pairs.as<int2>());
    CRISPY_LAUNCHED(ctx);
  }
  // no wait here: every consumer of the result is ordered on the same stream
  km.valid = true;
  if (count) *count = km.count;
  return CRISPY_OK;
}

int crispy_kmer_fetch(crispy_ctx *ctx, int64_t first, int64_t count, int32_t *pairs, float *dists) {
  if (!ctx) return CRISPY_ERR_ARG;
  KmerState &km = ctx->km;
  if (!km.valid) return set_err(ctx, CRISPY_ERR_STATE, "no k-mer result");
  if (first < 0 || count < 0 || first + count > km.count)
    return set_err(ctx, CRISPY_ERR_ARG, "fetch range outside the result");
  if (count == 0) return CRISPY_OK;
  CRISPY_HIP(ctx, hipSetDevice(ctx->device));
  if (pairs)
    CRISPY_HIP(ctx, hipMemcpyAsync(pairs, km.pairs.as<int2>() + first, sizeof(int2) * count,
                                   hipMemcpyDeviceToHost, ctx->stream));
  if (dists)
    CRISPY_HIP(ctx, hipMemcpyAsync(dists, km.dists_alt.as<float>() + first, sizeof(float) * count,
                                   hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

}  // extern "C"

// Used by the NW driver for the fused path.
namespace crispy {
const int2 *kmer_result_pairs(crispy_ctx *ctx) { return ctx->km.pairs.as<int2>(); }
}  // namespace crispy
