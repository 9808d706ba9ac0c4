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
//     Pairs are bucketed by kB (template) and sorted by (L2, L1) so the
//     64 lanes of a wave walk rows in lock step.  Reads are 2-bit packed; a
//     row's match bits come from XOR of a sliding 2-bit window of read 1
//     with read 2's base replicated.  Requires ACGT-only reads.
//   nw_generic_kernel  one lane per pair, the row state in a lane-
//     interleaved global scratch, raw byte compares.  Any kB, full mode and
//     reads with other bytes (N, lowercase, ...).
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.hpp"

namespace crispy {

const int2 *kmer_result_pairs(crispy_ctx *ctx);

constexpr uint32_t kGenericKey = 0xFFFFFFFFu;

// ---------------------------------------------------------------- bucketing
// Sort key per pair: (kB << 22 | L2 << 11 | L1), lengths clamped to 2047,
// for the register and sliced kernels; all ones for the generic kernel.  Per-bucket pair and cell counts are
// reduced in LDS first (a handful of buckets would serialise global atomics).
__global__ __launch_bounds__(256) void nw_bucket_kernel(
    const int2 *__restrict__ pairs, int64_t np, const int32_t *__restrict__ len,
    const uint8_t *__restrict__ clean, int band, uint32_t *__restrict__ keys,
    uint32_t *__restrict__ idx, unsigned int *__restrict__ hist,
    unsigned long long *__restrict__ cells) {
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
    uint32_t key = kGenericKey;
    int bucket = kGenericBucket;
    int64_t nc;
    if (band >= 2) {
      const int kb = (L1 + L2) / (2 * band);
      nc = band_cells(L1, L2, kb);
      if (kb <= kKbSliceMax && clean[pr.x] && clean[pr.y]) {
        key = ((uint32_t)kb << 22) | ((uint32_t)min(L2, 2047) << 11) | (uint32_t)min(L1, 2047);
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

// ------------------------------------------------------- register kernel
// Per band position d the state of the previous (then current) row:
//   Mt = M + 4i               (frame with the -4 mismatch folded in: diag
//                              becomes Mt(i-1,j-1) + 9*match)
//   Mv = Mt + 4 + pen_v(U)    (value seen as "left" by row i+1)
//   LA = (N << 16) | F        N = AL - ML, F = AL - j: a diagonal step adds
//                              match << 16, up adds nothing, left adds 1.
enum { kRowFull = 0, kRowCapture = 1, kRowTop = 2 };

template <int KB, int MODE>
__device__ __forceinline__ void nw_row(const int i, int (&Mt)[2 * KB + 1], int (&Mv)[2 * KB + 1],
                                       uint32_t (&LA)[2 * KB + 1],
                                       const uint32_t (&m)[(2 * KB + 1 + 15) / 16], const int dl,
                                       int &capM, uint32_t &capLA) {
  constexpr int W = 2 * KB + 1;
  int Mh = kNeg;      // "up" value of cell d-1 of this row (with pen_h applied)
  uint32_t LAu = 0u;  // its LA
#pragma unroll
  for (int d = 0; d < W; ++d) {
    if (MODE == kRowTop) {
      if (d < KB - i) continue;  // column j = i-kB+d < 0: not part of the matrix
      if (d == KB - i) {         // boundary cell (i, 0): M 0, ML 0, AL i
        Mt[d] = 4 * i;
        Mv[d] = 4 * i + 4;
        LA[d] = ((uint32_t)i << 16) | (uint32_t)i;
        Mh = 4 * i;
        LAu = LA[d];
        continue;
      }
    }
    const uint32_t s = (m[d >> 4] >> (2 * (d & 15))) & 1u;
    const int diag = Mt[d] + (int)(9u * s);
    const int left = d < W - 1 ? Mv[d + 1] : kNeg;
    const int up = d > 0 ? Mh : kNeg;
    const int M = max(max(diag, up), left);
    const bool cd = M == diag;
    const bool cu = M == up;
    const uint32_t lad = LA[d] + (s << 16);
    const uint32_t lal = d < W - 1 ? LA[d + 1] + 1u : 0u;
    const uint32_t la = cd ? lad : (cu ? LAu : lal);
    // pen_h (low half) and 4 + pen_v (high half) of the winner in one select;
    // the sign-extending half reads fold into SDWA operands of the adds.
    constexpr uint32_t kPenDiag = ((uint32_t)(4 + kGO) << 16) | (uint16_t)kGO;
    constexpr uint32_t kPenUp = ((uint32_t)4 << 16) | (uint16_t)kGE;
    constexpr uint32_t kPenLeft = ((uint32_t)(4 + kGE) << 16) | 0u;
    const uint32_t pk = cd ? kPenDiag : (cu ? kPenUp : kPenLeft);
    Mh = M + (int)(int16_t)(pk & 0xFFFFu);
    LAu = la;
    Mt[d] = M;
    Mv[d] = M + ((int)pk >> 16);
    // Materialise Mv here: left to itself the compiler sinks this select
    // into the next row and keeps both comparison masks of every cell alive
    // (SGPR pairs, spilled to VGPR lanes).
    asm volatile("" : "+v"(Mv[d]));
    LA[d] = la;
    if (MODE != kRowFull) {
      const bool c = d == dl;
      capM = c ? M : capM;
      capLA = c ? la : capLA;
    }
  }
}

__device__ __forceinline__ uint32_t load_word(const uint32_t *s, int w, int nw) {
  return (w >= 0 && w < nw) ? s[w] : 0u;
}

// 16 bases starting at base index q (may be negative or past the end: 0s).
__device__ __forceinline__ uint32_t get16(const uint32_t *s, int q, int nw) {
  const int w = q >> 4;
  const int o = q & 15;
  const uint32_t lo = load_word(s, w, nw);
  const uint32_t hi = load_word(s, w + 1, nw);
  return o ? ((lo >> (2 * o)) | (hi << (32 - 2 * o))) : lo;
}

// Occupancy target: the row state is 3 VGPRs per band cell plus ~64 others;
// asking for the waves that this fits keeps the scheduler from hoisting the
// chain-independent half of every cell (which would double the live state).
template <int KB>
constexpr int nw_reg_waves() {
  return (3 * (2 * KB + 1) + 64) <= 128 ? 4 : (3 * (2 * KB + 1) + 64) <= 168 ? 3 : 2;
}

template <int KB>
__global__ __launch_bounds__(64, nw_reg_waves<KB>()) void nw_band_reg_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, int32_t count, float *__restrict__ out) {
  constexpr int W = 2 * KB + 1;
  constexpr int NMW = (W + 15) / 16;  // match words per row
  constexpr int NB = NMW + 1;         // window words (one of lookahead)
  const int lane = threadIdx.x;
  const int base = blockIdx.x * 64;
  const int t = base + lane;
  const bool active = t < count;
  const uint32_t pidx = order[active ? t : count - 1];
  const int2 pr = pairs[pidx];
  const int L1 = len[pr.x], L2 = len[pr.y];
  const uint32_t *s1 = pack + poff[pr.x];
  const uint32_t *s2 = pack + poff[pr.y];
  const int nw1 = (L1 + 15) >> 4, nw2 = (L2 + 15) >> 4;

  // wave-uniform row bounds (pairs are sorted by (kB, L2, L1), so these are close)
  int l1m = L1, l2lo = L2, l2hi = L2;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    l1m = min(l1m, __shfl_xor(l1m, off, 64));
    l2lo = min(l2lo, __shfl_xor(l2lo, off, 64));
    l2hi = max(l2hi, __shfl_xor(l2hi, off, 64));
  }
  const int L1min = __builtin_amdgcn_readfirstlane(l1m);
  const int L2min = __builtin_amdgcn_readfirstlane(l2lo);
  const int L2max = __builtin_amdgcn_readfirstlane(l2hi);

  int Mt[W], Mv[W];
  uint32_t LA[W];
#pragma unroll
  for (int d = 0; d < W; ++d) {  // row 0: (0, j) = M 0, ML 0, AL j for 0 <= j <= kB
    const int j = d - KB;
    Mt[d] = j >= 0 ? 0 : kNeg;
    Mv[d] = j >= 0 ? 4 : kNeg;
    LA[d] = j >= 0 ? ((uint32_t)j << 16) : 0u;
  }
  // window of read-1 bases: position d of row i is base j-1 = i-kB-1+d
  uint32_t B[NB];
  int p0 = -KB;
#pragma unroll
  for (int k = 0; k < NB; ++k) B[k] = get16(s1, p0 + 16 * k, nw1);
  int qn = p0 + 16 * NB;
  uint32_t F = get16(s1, qn, nw1);
  int fcnt = 16;
  uint32_t R2 = 0;

  int bestI = kNeg, bestJ = kNeg;  // last row (i = L2) / last column (j = L1)
  uint32_t bestILA = 0, bestJLA = 0;
  int bestIj = 0;

  // One row of the DP, the end-cell bookkeeping, and the window slide.
  auto step = [&](const int i, auto mode_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    if (((i - 1) & 15) == 0) R2 = load_word(s2, (i - 1) >> 4, nw2);
    const uint32_t bb = ((R2 >> (2 * ((i - 1) & 15))) & 3u) * 0x55555555u;
    uint32_t m[NMW];
#pragma unroll
    for (int k = 0; k < NMW; ++k) {
      const uint32_t x = B[k] ^ bb;
      m[k] = ~(x | (x >> 1)) & 0x55555555u;
    }
    const int dl = L1 - i + KB;  // position of column L1 in this row
    int capM = kNeg;
    uint32_t capLA = 0;
    nw_row<KB, MODE>(i, Mt, Mv, LA, m, dl, capM, capLA);
    if (MODE != kRowFull && dl >= 0 && dl < W && i <= L2) {
      const int Mr = capM - 4 * i;
      if (Mr >= bestJ) {
        bestJ = Mr;
        bestJLA = capLA;
      }
    }
    if (i >= L2min && i == L2) {
#pragma unroll
      for (int d = 0; d < W; ++d) {
        const int j = i - KB + d;
        if (j >= 1 && j <= L1) {
          const int Mr = Mt[d] - 4 * i;
          if (Mr >= bestI) {
            bestI = Mr;
            bestILA = LA[d];
            bestIj = j;
          }
        }
      }
    }
    // slide the window by one base
#pragma unroll
    for (int k = 0; k < NB - 1; ++k) B[k] = __builtin_amdgcn_alignbit(B[k + 1], B[k], 2);
    B[NB - 1] = __builtin_amdgcn_alignbit(F, B[NB - 1], 2);
    F >>= 2;
    if (--fcnt == 0) {
      qn += 16;
      F = get16(s1, qn, nw1);
      fcnt = 16;
    }
  };
  // Rows 1..kB hold the boundary column; rows below L1min-kB cannot reach
  // column L1 of any lane.  Three loops, one row variant each, so the
  // state keeps one register assignment per loop.
  const int top_end = min(KB, L2max);
  const int mid_end = min(L2max, max(top_end, L1min - KB - 1));
  int i = 1;
  for (; i <= top_end; ++i) step(i, std::integral_constant<int, kRowTop>{});
  for (; i <= mid_end; ++i) step(i, std::integral_constant<int, kRowFull>{});
  for (; i <= L2max; ++i) step(i, std::integral_constant<int, kRowCapture>{});
  uint32_t la;
  int j;
  if (bestI > bestJ) {
    la = bestILA;
    j = bestIj;
  } else {
    la = bestJLA;
    j = L1;
  }
  const int AL = (int)(la & 0xFFFFu) + j;
  const int ML = AL - (int)(la >> 16);
  if (active) out[pidx] = (float)ML / (float)AL;
}

// -------------------------------------------------------- sliced kernel
// Wide bands (kB > kKbRegMax): G = ceil(W / 24) lanes per pair, lane g owns
// band positions [24g, 24g + 24) of every row in VGPRs (same per-cell state
// and arithmetic as nw_band_reg_kernel).  Lane g processes row t - g at step
// t, one row behind lane g - 1, so per step it needs only
//   * "up" for its first cell: lane g-1's last cell of the same row, sent at
//     the end of the previous step, and
//   * "left" for its last cell: lane g+1's first cell of the previous row,
//     which lane g+1 computes first in this step and sends mid-step.
// Positions >= W in the last lane are dead; the last real cell (KEND in the
// last lane, KEND = (W - 1) mod 24) sees an out-of-band left neighbour.
template <int KEND, int MODE>
__device__ __forceinline__ void nw_slice_cell(const int k, const int i, const uint32_t s,
                                              int &Mt_k, int &Mv_k, uint32_t &LA_k,
                                              const int leftM, const uint32_t leftLA, int &Mh,
                                              uint32_t &LAu, const int k0, const int kc,
                                              int &capM, uint32_t &capLA) {
  const int diag = Mt_k + (int)(9u * s);
  int M = max(max(diag, Mh), leftM);
  const bool cd = M == diag;
  const bool cu = M == Mh;
  uint32_t la = cd ? LA_k + (s << 16) : (cu ? LAu : leftLA + 1u);
  constexpr uint32_t kPenDiag = ((uint32_t)(4 + kGO) << 16) | (uint16_t)kGO;
  constexpr uint32_t kPenUp = ((uint32_t)4 << 16) | (uint16_t)kGE;
  constexpr uint32_t kPenLeft = ((uint32_t)(4 + kGE) << 16) | 0u;
  const uint32_t pk = cd ? kPenDiag : (cu ? kPenUp : kPenLeft);
  int mh = M + (int)(int16_t)(pk & 0xFFFFu);
  int mv = M + ((int)pk >> 16);
  if (MODE == kRowTop) {  // boundary cell (i, 0): M 0, ML 0, AL i
    const bool b = k == k0;
    M = b ? 4 * i : M;
    mv = b ? 4 * i + 4 : mv;
    la = b ? (((uint32_t)i << 16) | (uint32_t)i) : la;
    mh = b ? 4 * i : mh;
  }
  Mt_k = M;
  Mv_k = mv;
  asm volatile("" : "+v"(Mv_k));
  LA_k = la;
  Mh = mh;
  LAu = la;
  if (MODE != kRowFull) {
    const bool c = k == kc;
    capM = c ? M : capM;
    capLA = c ? la : capLA;
  }
}

constexpr int kSliceCells = 24;

template <int KEND>
__global__ __launch_bounds__(64) void nw_band_sliced_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, int32_t count, int KB, int G, float *__restrict__ out) {
  constexpr int C = kSliceCells;
  constexpr int NMW = (C + 15) / 16;
  constexpr int NB = NMW + 1;
  __shared__ int red[64][6];
  const int W = 2 * KB + 1;
  const int lane = threadIdx.x;
  const int groups = 64 / G;
  const int grp = lane / G;
  const int g = lane - grp * G;
  const int slot = blockIdx.x * groups + grp;
  const bool has_pair = grp < groups && slot < count;
  const uint32_t pidx = order[has_pair ? slot : count - 1];
  const int2 pr = pairs[pidx];
  const int L1 = len[pr.x];
  const int L2 = has_pair ? len[pr.y] : 0;
  const uint32_t *s1 = pack + poff[pr.x];
  const uint32_t *s2 = pack + poff[pr.y];
  const int nw1 = (L1 + 15) >> 4;
  const bool is_end = g == G - 1;
  const int src_up = g > 0 ? lane - 1 : lane;
  const int src_left = g < G - 1 ? lane + 1 : lane;
  const int base_d = g * C;

  int tm = has_pair ? L2 + g : 0;
  int l1m = has_pair ? L1 : 0x7fffffff;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    tm = max(tm, __shfl_xor(tm, off, 64));
    l1m = min(l1m, __shfl_xor(l1m, off, 64));
  }
  const int tmax = __builtin_amdgcn_readfirstlane(tm);
  const int L1min = __builtin_amdgcn_readfirstlane(l1m);

  int Mt[C], Mv[C];
  uint32_t LA[C];
#pragma unroll
  for (int k = 0; k < C; ++k) {  // row 0 within the band; dead cells out of band
    const int j = base_d + k - KB;
    const bool live = !(is_end && k > KEND) && j >= 0;
    Mt[k] = live ? 0 : kNeg;
    Mv[k] = live ? 4 : kNeg;
    LA[k] = live ? ((uint32_t)j << 16) : 0u;
  }
  // read-1 window: cell k of lane g at step t is base t - g - kB - 1 + 24g + k
  const int p0 = base_d - g - KB;
  uint32_t B[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) B[k] = get16(s1, p0 + 16 * k, nw1);
  int qn = p0 + 16 * NB;
  uint32_t F = get16(s1, qn, nw1);
  int fcnt = 16;

  int upM = kNeg;
  uint32_t upLA = 0;
  int bIM = kNeg, bIj = 0, bJM = kNeg, bJr = 0;
  uint32_t bILA = 0, bJLA = 0;

  auto step = [&](const int t, auto mode_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    const int i = t - g;
    const bool active = has_pair && i >= 1 && i <= L2;
    const int im1 = i - 1;
    const uint32_t w2 = (im1 >= 0 && im1 < L2) ? s2[im1 >> 4] : 0u;
    const uint32_t bb = ((w2 >> (2 * (im1 & 15))) & 3u) * 0x55555555u;
    uint32_t m[NMW];
#pragma unroll
    for (int k = 0; k < NMW; ++k) {
      const uint32_t x = B[k] ^ bb;
      m[k] = ~(x | (x >> 1)) & 0x55555555u;
    }
    const int k0 = KB - i - base_d;      // the boundary cell (i, 0) in this lane
    const int kc = L1 - i + KB - base_d;  // the last-column cell (i, L1) in this lane
    int capM = kNeg;
    uint32_t capLA = 0;
    int Mh = upM;
    uint32_t LAu = upLA;
    if (active) {
      const bool last0 = KEND == 0 && is_end;
      nw_slice_cell<KEND, MODE>(0, i, m[0] & 1u, Mt[0], Mv[0], LA[0], last0 ? kNeg : Mv[1], LA[1],
                                Mh, LAu, k0, kc, capM, capLA);
    }
    // lane g+1's first cell of row i-1 is this lane's "left" for its last cell
    const int xM = __shfl(Mv[0], src_left, 64);
    const uint32_t xLA = (uint32_t)__shfl((int)LA[0], src_left, 64);
    if (active) {
#pragma unroll
      for (int k = 1; k < C; ++k) {
        const uint32_t s = (m[k >> 4] >> (2 * (k & 15))) & 1u;
        int lM;
        uint32_t lLA;
        if (k == C - 1) {
          lM = is_end ? kNeg : xM;
          lLA = xLA;
        } else {
          lM = (k == KEND && is_end) ? kNeg : Mv[k + 1];
          lLA = LA[k + 1];
        }
        nw_slice_cell<KEND, MODE>(k, i, s, Mt[k], Mv[k], LA[k], lM, lLA, Mh, LAu, k0, kc, capM,
                                  capLA);
      }
      if (MODE != kRowFull && kc >= 0 && kc < C && i >= L1 - KB) {
        const int Mr = capM - 4 * i;
        if (Mr >= bJM) {
          bJM = Mr;
          bJr = i;
          bJLA = capLA;
        }
      }
      if (i == L2) {
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const int j = i - KB + base_d + k;
          if (j >= 1 && j <= L1 && base_d + k < W) {
            const int Mr = Mt[k] - 4 * i;
            if (Mr >= bIM) {
              bIM = Mr;
              bIj = j;
              bILA = LA[k];
            }
          }
        }
      }
    }
    // this lane's last cell of row i is lane g+1's "up" at the next step
    const int nM = __shfl(Mh, src_up, 64);
    const uint32_t nLA = (uint32_t)__shfl((int)LAu, src_up, 64);
    upM = g > 0 ? nM : kNeg;
    upLA = nLA;
#pragma unroll
    for (int k = 0; k < NB - 1; ++k) B[k] = __builtin_amdgcn_alignbit(B[k + 1], B[k], 2);
    B[NB - 1] = __builtin_amdgcn_alignbit(F, B[NB - 1], 2);
    F >>= 2;
    if (--fcnt == 0) {
      qn += 16;
      F = get16(s1, qn, nw1);
      fcnt = 16;
    }
  };
  const int top_end = min(KB + G - 1, tmax);
  const int mid_end = max(top_end, min(tmax, L1min - KB - 1));
  int t = 1;
  for (; t <= top_end; ++t) step(t, std::integral_constant<int, kRowTop>{});
  for (; t <= mid_end; ++t) step(t, std::integral_constant<int, kRowFull>{});
  for (; t <= tmax; ++t) step(t, std::integral_constant<int, kRowCapture>{});

  red[lane][0] = bIM;
  red[lane][1] = bIj;
  red[lane][2] = (int)bILA;
  red[lane][3] = bJM;
  red[lane][4] = bJr;
  red[lane][5] = (int)bJLA;
  __syncthreads();
  if (has_pair && g == 0) {
    int iM = kNeg, ij = 0, jM = kNeg, jr = 0;
    uint32_t iLA = 0, jLA = 0;
    for (int q = lane; q < lane + G; ++q) {  // last maximum: larger j (row) wins ties
      if (red[q][0] > iM || (red[q][0] == iM && red[q][1] > ij)) {
        iM = red[q][0];
        ij = red[q][1];
        iLA = (uint32_t)red[q][2];
      }
      if (red[q][3] > jM || (red[q][3] == jM && red[q][4] > jr)) {
        jM = red[q][3];
        jr = red[q][4];
        jLA = (uint32_t)red[q][5];
      }
    }
    uint32_t la;
    int j;
    if (iM > jM) {
      la = iLA;
      j = ij;
    } else {
      la = jLA;
      j = L1;
    }
    const int AL = (int)(la & 0xFFFFu) + j;
    const int ML = AL - (int)(la >> 16);
    out[pidx] = (float)ML / (float)AL;
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
                               uint8_t *__restrict__ flags) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x)
    flags[p] = nw_passes(dist[p], thr) ? 1 : 0;
}

__global__ void nw_gather_kernel(const uint32_t *__restrict__ sel, const unsigned int *__restrict__ nsel,
                                 const int2 *__restrict__ pairs, const float *__restrict__ dist,
                                 int2 *__restrict__ out_pairs, float *__restrict__ out_dists) {
  const int64_t n = *nsel;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = sel[k];
    out_pairs[k] = pairs[p];
    out_dists[k] = dist[p];
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

static inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// ------------------------------------------------------------ host driver
template <int KB>
static void launch_reg(hipStream_t s, const ReadStore &rs, const int2 *pairs, const uint32_t *order,
                       int32_t count, float *out) {
  nw_band_reg_kernel<KB><<<(count + 63) / 64, 64, 0, s>>>(
      rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(), pairs, order, count, out);
}

using RegFn = void (*)(hipStream_t, const ReadStore &, const int2 *, const uint32_t *, int32_t, float *);

template <int... KBs>
struct RegTable {
  static constexpr RegFn fns[sizeof...(KBs)] = {launch_reg<KBs>...};
};
template <int... KBs>
constexpr RegFn RegTable<KBs...>::fns[sizeof...(KBs)];

template <int N, int... KBs>
struct MakeTable : MakeTable<N - 1, N - 1, KBs...> {};
template <int... KBs>
struct MakeTable<0, KBs...> {
  using type = RegTable<KBs...>;
};
using RegFns = MakeTable<kKbRegMax + 1>::type;

template <int KEND>
static void launch_sliced(hipStream_t s, const ReadStore &rs, const int2 *pairs, const uint32_t *order,
                          int32_t count, int kb, float *out) {
  const int W = 2 * kb + 1;
  const int G = (W + kSliceCells - 1) / kSliceCells;
  const int groups = 64 / G;
  nw_band_sliced_kernel<KEND><<<(count + groups - 1) / groups, 64, 0, s>>>(
      rs.pack.as<uint32_t>(), rs.poff.as<int64_t>(), rs.len.as<int32_t>(), pairs, order, count, kb, G,
      out);
}

using SliceFn = void (*)(hipStream_t, const ReadStore &, const int2 *, const uint32_t *, int32_t, int,
                         float *);
// KEND = (W - 1) mod 24 = 2kB mod 24 is even
static const SliceFn kSliceFns[kSliceCells / 2] = {
    launch_sliced<0>,  launch_sliced<2>,  launch_sliced<4>,  launch_sliced<6>,
    launch_sliced<8>,  launch_sliced<10>, launch_sliced<12>, launch_sliced<14>,
    launch_sliced<16>, launch_sliced<18>, launch_sliced<20>, launch_sliced<22>};

// Distances of np device pairs into nw.dist (indexed like the pairs).
static int nw_core(crispy_ctx *ctx, const int2 *d_pairs, int64_t np, int band) {
  ReadStore &rs = ctx->reads[CRISPY_READS_NW];
  NwState &nw = ctx->nw;
  if (np > 0xFFFFFFFFll) return set_err(ctx, CRISPY_ERR_ARG, "too many pairs in one call");
  CRISPY_TRY(ensure(ctx, nw.keys, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.keys_alt, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.idx, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.idx_alt, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.dist, sizeof(float) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.hist, sizeof(unsigned int) * kNumBuckets));
  CRISPY_TRY(ensure(ctx, nw.cells, sizeof(unsigned long long) * kNumBuckets));
  if (np == 0) return CRISPY_OK;
  CRISPY_HIP(ctx, hipMemsetAsync(nw.hist.p, 0, sizeof(unsigned int) * kNumBuckets, ctx->stream));
  CRISPY_HIP(ctx, hipMemsetAsync(nw.cells.p, 0, sizeof(unsigned long long) * kNumBuckets, ctx->stream));
  hipEvent_t ev;
  timer_begin(ctx, "nw_bucket", &ev);
  nw_bucket_kernel<<<(int)std::min<int64_t>((np + 255) / 256, 1024), 256, 0, ctx->stream>>>(
      d_pairs, np, rs.len.as<int32_t>(), rs.clean.as<uint8_t>(), band, nw.keys.as<uint32_t>(),
      nw.idx.as<uint32_t>(), nw.hist.as<unsigned int>(), nw.cells.as<unsigned long long>());
  CRISPY_LAUNCHED(ctx);
  size_t tbytes = 0;
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tbytes, nw.keys.as<uint32_t>(),
                                            nw.keys_alt.as<uint32_t>(), nw.idx.as<uint32_t>(),
                                            nw.idx_alt.as<uint32_t>(), (size_t)np, 0, 32, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(ctx->temp.p, tbytes, nw.keys.as<uint32_t>(),
                                            nw.keys_alt.as<uint32_t>(), nw.idx.as<uint32_t>(),
                                            nw.idx_alt.as<uint32_t>(), (size_t)np, 0, 32, ctx->stream));
  timer_end(ctx, "nw_bucket", ev, (double)np);
  unsigned int hist[kNumBuckets];
  unsigned long long cells[kNumBuckets];
  CRISPY_HIP(ctx, hipMemcpyAsync(hist, nw.hist.p, sizeof hist, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(cells, nw.cells.p, sizeof cells, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint32_t *order = nw.idx_alt.as<uint32_t>();
  float *out = nw.dist.as<float>();
  // Buckets are independent: fan them out over the worker streams, largest
  // first, so small buckets do not leave the chip idle between launches.
  int64_t start[kNumBuckets];
  int64_t pos = 0;
  std::vector<int> todo;
  double total_cells = 0;
  for (int b = 0; b < kNumBuckets; ++b) {
    start[b] = pos;
    pos += hist[b];
    if (hist[b]) todo.push_back(b);
    total_cells += (double)cells[b];
  }
  std::sort(todo.begin(), todo.end(), [&](int a, int b) { return cells[a] > cells[b]; });
  const int64_t ng = hist[kGenericBucket];
  const int gblocks = (int)((std::min<int64_t>(ng, 32768) + 255) / 256);
  if (ng > 0)  // before the fork: ensure() may free and reallocate
    CRISPY_TRY(ensure(ctx, nw.scratch, sizeof(int2) * (size_t)gblocks * 256 * (rs.maxlen + 1)));
  hipEvent_t phase;
  timer_begin(ctx, "nw_dp_phase", &phase);
  CRISPY_TRY(fork_workers(ctx));
  int wk = 0;
  for (int b : todo) {
    hipStream_t s = ctx->workers[wk++ % kNumWorkers];
    if (b == kGenericBucket) {
      timer_begin(ctx, "nw_generic", &ev, s);
      nw_generic_kernel<<<gblocks, 256, 0, s>>>(rs.bases.as<uint8_t>(), rs.off.as<int64_t>(),
                                               rs.len.as<int32_t>(), d_pairs, order + start[b],
                                               (int32_t)ng, band, nw.scratch.as<int2>(), out);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "nw_generic", ev, (double)cells[b], s);
    } else if (b <= kKbRegMax) {
      timer_begin(ctx, "nw_band_reg", &ev, s);
      RegFns::fns[b](s, rs, d_pairs, order + start[b], (int32_t)hist[b], out);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "nw_band_reg", ev, (double)cells[b], s);
    } else {
      timer_begin(ctx, "nw_band_sliced", &ev, s);
      kSliceFns[((2 * b) % kSliceCells) / 2](s, rs, d_pairs, order + start[b], (int32_t)hist[b], b, out);
      CRISPY_LAUNCHED(ctx);
      timer_end(ctx, "nw_band_sliced", ev, (double)cells[b], s);
    }
  }
  CRISPY_TRY(join_workers(ctx));
  timer_end(ctx, "nw_dp_phase", phase, total_cells);
  return CRISPY_OK;
}

// Keeps the pairs whose distance passes, in input order.
static int nw_filter(crispy_ctx *ctx, const int2 *d_pairs, int64_t np, float thr, int64_t *count) {
  NwState &nw = ctx->nw;
  CRISPY_TRY(ensure(ctx, nw.flags, np + 1));
  CRISPY_TRY(ensure(ctx, nw.sel, sizeof(uint32_t) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.nsel, sizeof(unsigned int)));
  CRISPY_TRY(ensure(ctx, nw.out_pairs, sizeof(int2) * (np + 1)));
  CRISPY_TRY(ensure(ctx, nw.out_dists, sizeof(float) * (np + 1)));
  *count = 0;
  if (np == 0) return CRISPY_OK;
  hipEvent_t ev;
  timer_begin(ctx, "nw_filter", &ev);
  nw_flag_kernel<<<grid_for(np), 256, 0, ctx->stream>>>(nw.dist.as<float>(), np, thr, nw.flags.as<uint8_t>());
  CRISPY_LAUNCHED(ctx);
  size_t tbytes = 0;
  rocprim::counting_iterator<uint32_t> it(0);
  CRISPY_HIP(ctx, rocprim::select(nullptr, tbytes, it, nw.flags.as<uint8_t>(), nw.sel.as<uint32_t>(),
                                  nw.nsel.as<unsigned int>(), (size_t)np, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, tbytes));
  CRISPY_HIP(ctx, rocprim::select(ctx->temp.p, tbytes, it, nw.flags.as<uint8_t>(), nw.sel.as<uint32_t>(),
                                  nw.nsel.as<unsigned int>(), (size_t)np, ctx->stream));
  nw_gather_kernel<<<grid_for(np), 256, 0, ctx->stream>>>(nw.sel.as<uint32_t>(), nw.nsel.as<unsigned int>(),
                                                           d_pairs, nw.dist.as<float>(),
                                                           nw.out_pairs.as<int2>(), nw.out_dists.as<float>());
  CRISPY_LAUNCHED(ctx);
  timer_end(ctx, "nw_filter", ev, (double)np);
  unsigned int c = 0;
  CRISPY_HIP(ctx, hipMemcpyAsync(&c, nw.nsel.p, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *count = c;
  return CRISPY_OK;
}

static int upload_pairs(crispy_ctx *ctx, const int32_t *pairs, int64_t np) {
  const int32_t n = ctx->reads[CRISPY_READS_NW].n;
  for (int64_t p = 0; p < 2 * np; ++p)
    if (pairs[p] < 0 || pairs[p] >= n)
      return set_err(ctx, CRISPY_ERR_ARG, "pair %lld names read %d outside [0, %d)", (long long)(p / 2),
                     pairs[p], n);
  CRISPY_TRY(ensure(ctx, ctx->nw.in_pairs, sizeof(int2) * (np + 1)));
  if (np > 0)
    CRISPY_HIP(ctx, hipMemcpyAsync(ctx->nw.in_pairs.p, pairs, sizeof(int2) * np, hipMemcpyHostToDevice,
                                   ctx->stream));
  return CRISPY_OK;
}

}  // namespace crispy

using namespace crispy;

extern "C" {

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
  CRISPY_TRY(nw_core(ctx, d, np, band));
  int64_t c = 0;
  CRISPY_TRY(nw_filter(ctx, d, np, threshold, &c));
  ctx->nw.count = c;
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
  nw.valid = false;
  // (i, j) keys in the out buffers, dist bits in keys/keys_alt
  CRISPY_TRY(ensure(ctx, nw.in_pairs, sizeof(int2) * n));
  CRISPY_TRY(ensure(ctx, nw.out_pairs, sizeof(uint64_t) * n));
  CRISPY_TRY(ensure(ctx, nw.scratch, sizeof(uint64_t) * n));
  CRISPY_TRY(ensure(ctx, nw.dist, sizeof(float) * n));
  CRISPY_TRY(ensure(ctx, nw.keys, sizeof(uint32_t) * n));
  CRISPY_TRY(ensure(ctx, nw.keys_alt, sizeof(uint32_t) * n));
  uint64_t *k0 = nw.out_pairs.as<uint64_t>(), *k1 = nw.scratch.as<uint64_t>();
  uint32_t *b0 = nw.keys.as<uint32_t>(), *b1 = nw.keys_alt.as<uint32_t>();
  CRISPY_HIP(ctx, hipMemcpyAsync(nw.in_pairs.p, pairs, sizeof(int2) * n, hipMemcpyHostToDevice, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(nw.dist.p, dists, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
  pack_pair_keys_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(nw.in_pairs.as<int2>(), n, k0);
  dist_bits_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(nw.dist.as<float>(), n, b0);
  CRISPY_LAUNCHED(ctx);
  // 1) by (i, j)   2) stable by dist  ->  (dist, i, j)
  size_t t1 = 0, t2 = 0;
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(nullptr, t1, k0, k1, b0, b1, (size_t)n, 0, 40, ctx->stream));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(nullptr, t2, b1, b0, k1, k0, (size_t)n, 0, 32, ctx->stream));
  CRISPY_TRY(ensure(ctx, ctx->temp, std::max(t1, t2)));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(ctx->temp.p, t1, k0, k1, b0, b1, (size_t)n, 0, 40, ctx->stream));
  CRISPY_HIP(ctx, rocprim::radix_sort_pairs(ctx->temp.p, t2, b1, b0, k1, k0, (size_t)n, 0, 32, ctx->stream));
  unpack_pair_keys_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(k0, b0, n, nw.in_pairs.as<int2>(),
                                                                 nw.dist.as<float>());
  CRISPY_LAUNCHED(ctx);
  CRISPY_HIP(ctx, hipMemcpyAsync(pairs, nw.in_pairs.p, sizeof(int2) * n, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipMemcpyAsync(dists, nw.dist.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  CRISPY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CRISPY_OK;
}

}  // extern "C"
