// Device bodies of the banded NW DP (register and lane-sliced forms) and the
// fused wave-task kernels built from them.  Included only by the per-family
// sources nw_fam_*.hip, so each family compiles in its own translation unit.
// Recurrence, state encoding and tie rules: see the head of nw.hip
// (genCPU.cpp:348-464).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"
#include "nw_fused.hpp"

namespace crispy {

// ------------------------------------------------------- register kernel
// Per band position d the state of the previous (then current) row:
//   Mt = M + 4i               (frame with the -4 mismatch folded in: diag
//                              becomes Mt(i-1,j-1) + 9*match)
//   Mv = Mt + 4 + pen_v(U)    (value seen as "left" by row i+1)
//   LA = (N << 16) | F        N = AL - ML, F = AL - j: a diagonal step adds
//                              match << 16, up adds nothing, left adds 1.
// Row variants: full (no end-cell work), capture (column L1 may be in band:
// record its cell), top (rows 1..kB: the boundary cell (i, 0) is in band),
// top with capture (both; only for reads shorter than about 2 kB).
enum { kRowFull = 0, kRowCapture = 1, kRowTop = 2, kRowTopCap = 3 };
constexpr bool row_top(int mode) { return mode == kRowTop || mode == kRowTopCap; }
constexpr bool row_captures(int mode) { return mode == kRowCapture || mode == kRowTopCap; }

// TR: the DP runs transposed (rows over the reference's read 1, columns over
// its read 2), so the reference's "up" move (same row) is this frame's
// "left" (previous row) and vice versa: of two equal gap moves the
// previous-row one wins, which keeps the reference's choice.
template <int KB, int MODE, bool TR>
__device__ __forceinline__ void nw_row(const int i, int (&Mt)[2 * KB + 1], int (&Mv)[2 * KB + 1],
                                       uint32_t (&LA)[2 * KB + 1],
                                       const uint32_t (&m)[(2 * KB + 1 + 15) / 16], const int dl,
                                       const int dmax, int &capM, uint32_t &capLA) {
  constexpr int W = 2 * KB + 1;
  int Mh = kNeg;      // "up" value of cell d-1 of this row (with pen_h applied)
  uint32_t LAu = 0u;  // its LA
#pragma unroll
  for (int d = 0; d < W; ++d) {
    // past column L1 of every lane: never read by a cell at or before column
    // L1 (the bound drops by one per row, like the cells' reach)
    if (row_captures(MODE) && d > dmax) continue;
    if (row_top(MODE)) {
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
    int M = max(max(diag, up), left);
    // opaque: otherwise the transposed frame's M == left becomes
    // left >= max(diag, up), which splits the max3 into two maxima
    asm volatile("" : "+v"(M));
    const bool cd = M == diag;
    const bool c2 = M == (TR ? left : up);  // the gap move that wins a tie
    const uint32_t lad = LA[d] + (s << 16);
    const uint32_t lal = d < W - 1 ? LA[d + 1] + 1u : 0u;
    const uint32_t la = cd ? lad : (c2 ? (TR ? lal : LAu) : (TR ? LAu : lal));
    // pen_h (low half) and 4 + pen_v (high half) of the winner in one select;
    // the sign-extending half reads fold into SDWA operands of the adds.
    constexpr uint32_t kPenDiag = ((uint32_t)(4 + kGO) << 16) | (uint16_t)kGO;
    constexpr uint32_t kPenUp = ((uint32_t)4 << 16) | (uint16_t)kGE;
    constexpr uint32_t kPenLeft = ((uint32_t)(4 + kGE) << 16) | 0u;
    const uint32_t pk = cd ? kPenDiag : (c2 ? (TR ? kPenLeft : kPenUp) : (TR ? kPenUp : kPenLeft));
    Mh = M + (int)(int16_t)(pk & 0xFFFFu);
    LAu = la;
    Mt[d] = M;
    Mv[d] = M + ((int)pk >> 16);
    // Materialise Mv here: left to itself the compiler sinks this select
    // into the next row and keeps both comparison masks of every cell alive
    // (SGPR pairs, spilled to VGPR lanes).
    asm volatile("" : "+v"(Mv[d]));
    LA[d] = la;
    if (row_captures(MODE)) {
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

// The DP of one wave's pairs: columns over s1 (L1 bases), rows over s2 (L2
// bases); TR when that is the transposed frame (s1 = the reference's read 2).
template <int KB, bool TR>
__device__ __forceinline__ void nw_reg_dp(const uint32_t *__restrict__ s1, const uint32_t *__restrict__ s2,
                                          const int L1, const int L2, const bool active,
                                          float *__restrict__ out) {
  constexpr int W = 2 * KB + 1;
  constexpr int NMW = (W + 15) / 16;  // match words per row
  constexpr int NB = NMW + 1;         // window words (one of lookahead)
  const int nw1 = (L1 + 15) >> 4, nw2 = (L2 + 15) >> 4;

  // Rows past L1 + kB hold no in-band cell: the last column is out of reach
  // and the last row, if that low, is out of band.  So a pair's DP ends at
  // row R = min(L2, L1 + kB).  Wave-uniform bounds (pairs are sorted by
  // (kB, R), so these are close); lanes past their own R sit out the rows
  // that the others still compute.
  const int R = min(L2, L1 + KB);
  int l1m = L1, l1M = L1, l2hi = R;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    l1m = min(l1m, __shfl_xor(l1m, off, 64));
    l1M = max(l1M, __shfl_xor(l1M, off, 64));
    l2hi = max(l2hi, __shfl_xor(l2hi, off, 64));
  }
  const int L1min = __builtin_amdgcn_readfirstlane(l1m);
  const int L1max = __builtin_amdgcn_readfirstlane(l1M);
  const int Rmax = __builtin_amdgcn_readfirstlane(l2hi);

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
    // rows past the pair's last in-band row R leave its registers alone, so
    // they still hold row R when the loop ends
    if (i <= R) nw_row<KB, MODE, TR>(i, Mt, Mv, LA, m, dl, L1max - i + KB, capM, capLA);
    if (row_captures(MODE) && dl >= 0 && dl < W && i <= L2) {
      const int Mr = capM - 4 * i;
      if (Mr >= bestJ) {
        bestJ = Mr;
        bestJLA = capLA;
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
  // Rows 1..kB hold the boundary column; rows above L1min-kB cannot reach
  // column L1 of any lane.  One loop per row variant, so the state keeps one
  // register assignment per loop.
  const int top_end = min(KB, Rmax);
  const int top_plain_end = min(top_end, L1min - KB - 1);
  const int mid_end = min(Rmax, max(top_end, L1min - KB - 1));
  int i = 1;
  for (; i <= top_plain_end; ++i) step(i, std::integral_constant<int, kRowTop>{});
  for (; i <= top_end; ++i) step(i, std::integral_constant<int, kRowTopCap>{});
  for (; i <= mid_end; ++i) step(i, std::integral_constant<int, kRowFull>{});
  for (; i <= Rmax; ++i) step(i, std::integral_constant<int, kRowCapture>{});
  // Last row: when R = L2 the registers hold row L2 (when R < L2, row L2 has
  // no in-band cell).
  if (R == L2) {
#pragma unroll
    for (int d = 0; d < W; ++d) {
      const int j = L2 - KB + d;
      if (j >= 1 && j <= L1) {
        const int Mr = Mt[d] - 4 * L2;
        if (Mr >= bestI) {
          bestI = Mr;
          bestILA = LA[d];
          bestIj = j;
        }
      }
    }
  }
  // the reference takes its last row's best only when it beats its last
  // column's (genCPU.cpp:461-464); transposed, those are this frame's last
  // column and last row
  uint32_t la;
  int j;
  if (TR ? !(bestJ > bestI) : bestI > bestJ) {
    la = bestILA;
    j = bestIj;
  } else {
    la = bestJLA;
    j = L1;
  }
  const int AL = (int)(la & 0xFFFFu) + j;
  const int ML = AL - (int)(la >> 16);
  if (active) *out = (float)ML / (float)AL;
}

// Rows over the shorter read: a pair's DP ends at row min(L2, L1 + kB), so
// with rows over the longer read up to kB rows run past column L1 with a
// shrinking band.  The recurrence is symmetric up to its tie rules, so a wave
// runs transposed (rows over read 1) when none of its pairs has read 1 longer
// than read 2, which is every pair of a length-sorted read set (i < j).
__device__ __forceinline__ bool wave_transposed(const bool lane_has_pair, const int La, const int Lb) {
  return __ballot(lane_has_pair && La > Lb) == 0;
}

// One wave: pairs order[base .. base+63] of a bucket of `count` pairs.
template <int KB>
__device__ __forceinline__ void nw_reg_wave(const uint32_t *__restrict__ pack,
                                            const int64_t *__restrict__ poff,
                                            const int32_t *__restrict__ len,
                                            const int2 *__restrict__ pairs,
                                            const uint32_t *__restrict__ order, int32_t count,
                                            int base, float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int t = base + lane;
  const bool active = t < count;
  const uint32_t pidx = order[active ? t : count - 1];
  const int2 pr = pairs[pidx];
  const int La = len[pr.x], Lb = len[pr.y];
  const uint32_t *sa = pack + poff[pr.x];
  const uint32_t *sb = pack + poff[pr.y];
  if (wave_transposed(active, La, Lb))
    nw_reg_dp<KB, true>(sb, sa, Lb, La, active, out + pidx);
  else
    nw_reg_dp<KB, false>(sa, sb, La, Lb, active, out + pidx);
}

// -------------------------------------------------------- sliced kernel
// DPP controls (GFX9 wave-wide shifts by one lane)
constexpr int kDppWaveShl1 = 0x130;  // lane l takes lane l+1
constexpr int kDppWaveShr1 = 0x138;  // lane l takes lane l-1

// Wide bands (kB > kKbRegMax): G = ceil(W / C) lanes per pair, lane g owns
// band positions [Cg, Cg + C) of every row in VGPRs (same per-cell state
// and arithmetic as nw_band_reg_kernel).  Lane g processes row t - g at step
// t, one row behind lane g - 1, so per step it needs only
//   * "up" for its first cell: lane g-1's last cell of the same row, sent at
//     the end of the previous step, and
//   * "left" for its last cell: lane g+1's first cell of the previous row,
//     which lane g+1 computes first in this step and sends mid-step.
// Positions >= W in the last lane are dead; the last real cell (KEND in the
// last lane, KEND = (W - 1) mod C) sees an out-of-band left neighbour.
template <int KEND, int MODE, bool TR>
__device__ __forceinline__ void nw_slice_cell(const int k, const int i, const uint32_t s,
                                              int &Mt_k, int &Mv_k, uint32_t &LA_k,
                                              const int leftM, const uint32_t leftLA, int &Mh,
                                              uint32_t &LAu, const int k0, const int kc,
                                              int &capM, uint32_t &capLA) {
  const int diag = Mt_k + (int)(9u * s);
  int M = max(max(diag, Mh), leftM);
  asm volatile("" : "+v"(M));  // keeps the max3 whole (see nw_row)
  const bool cd = M == diag;
  const bool c2 = M == (TR ? leftM : Mh);  // the gap move that wins a tie (see nw_row)
  uint32_t la = cd ? LA_k + (s << 16) : (c2 ? (TR ? leftLA + 1u : LAu) : (TR ? LAu : leftLA + 1u));
  constexpr uint32_t kPenDiag = ((uint32_t)(4 + kGO) << 16) | (uint16_t)kGO;
  constexpr uint32_t kPenUp = ((uint32_t)4 << 16) | (uint16_t)kGE;
  constexpr uint32_t kPenLeft = ((uint32_t)(4 + kGE) << 16) | 0u;
  uint32_t pk = cd ? kPenDiag : (c2 ? (TR ? kPenLeft : kPenUp) : (TR ? kPenUp : kPenLeft));
  if (row_top(MODE)) {  // boundary cell (i, 0): M 0, ML 0, AL i; pen_h = pen_v = 0
    constexpr uint32_t kPenBoundary = (uint32_t)4 << 16;
    const bool b = k == k0;
    M = b ? 4 * i : M;
    la = b ? (((uint32_t)i << 16) | (uint32_t)i) : la;
    pk = b ? kPenBoundary : pk;
  }
  const int mh = M + (int)(int16_t)(pk & 0xFFFFu);
  const int mv = M + ((int)pk >> 16);
  Mt_k = M;
  Mv_k = mv;
  asm volatile("" : "+v"(Mv_k));
  LA_k = la;
  Mh = mh;
  LAu = la;
  if (row_captures(MODE)) {
    const bool c = k == kc;
    capM = c ? M : capM;
    capLA = c ? la : capLA;
  }
}

// The DP of one wave's 64 / G pairs (lane g of group grp; columns over s1,
// rows over s2, L2 = 0 for a lane without a pair); TR as in nw_reg_dp.  `red`
// is the wave's 64 x 6 LDS reduction buffer.
template <int C, int KEND, bool TR>
__device__ __forceinline__ void nw_sliced_dp(const uint32_t *__restrict__ s1, const uint32_t *__restrict__ s2,
                                             const int L1, const int L2, const bool has_pair, const int KB,
                                             const int G, const int g, int (*red)[6],
                                             float *__restrict__ out) {
  constexpr int NMW = (C + 15) / 16;
  constexpr int NB = NMW + 1;
  const int W = 2 * KB + 1;
  const int lane = threadIdx.x & 63;
  const int nw1 = (L1 + 15) >> 4;
  const bool is_end = g == G - 1;
  const int base_d = g * C;

  // the pair's DP ends at row R = min(L2, L1 + kB) (no in-band cell below)
  const int R = min(L2, L1 + KB);
  int tm = has_pair ? R + g : 0;
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
  // read-2 words: R2c holds rows 16w+1 .. 16w+16 of the current row i
  // (w = (i-1) >> 4), R2n the next 16, loaded 16 steps before it is needed
  const int nw2 = (L2 + 15) >> 4;
  uint32_t R2c = load_word(s2, 0, nw2), R2n = load_word(s2, 1, nw2);

  auto step = [&](const int t, auto mode_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    const int i = t - g;
    const bool active = has_pair && i >= 1 && i <= R;
    const int im1 = i - 1;
    if (im1 > 0 && (im1 & 15) == 0) {
      R2c = R2n;
      R2n = load_word(s2, (im1 >> 4) + 1, nw2);
    }
    const uint32_t w2 = (im1 >= 0 && im1 < L2) ? R2c : 0u;
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
      nw_slice_cell<KEND, MODE, TR>(0, i, m[0] & 1u, Mt[0], Mv[0], LA[0], last0 ? kNeg : Mv[1], LA[1],
                                    Mh, LAu, k0, kc, capM, capLA);
    }
    // lane g+1's first cell of row i-1 is this lane's "left" for its last cell
    // (a DPP wave shift: lane l reads lane l+1; the last lane of a group ignores it)
    const int xM = __builtin_amdgcn_update_dpp(Mv[0], Mv[0], kDppWaveShl1, 0xF, 0xF, false);
    const uint32_t xLA = (uint32_t)__builtin_amdgcn_update_dpp((int)LA[0], (int)LA[0], kDppWaveShl1, 0xF,
                                                               0xF, false);
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
        nw_slice_cell<KEND, MODE, TR>(k, i, s, Mt[k], Mv[k], LA[k], lM, lLA, Mh, LAu, k0, kc, capM,
                                      capLA);
      }
      if (row_captures(MODE) && kc >= 0 && kc < C && i >= L1 - KB) {
        const int Mr = capM - 4 * i;
        if (Mr >= bJM) {
          bJM = Mr;
          bJr = i;
          bJLA = capLA;
        }
      }
    }
    // this lane's last cell of row i is lane g+1's "up" at the next step
    // (a DPP wave shift: lane l reads lane l-1; the first lane of a group ignores it)
    const int nM = __builtin_amdgcn_update_dpp(Mh, Mh, kDppWaveShr1, 0xF, 0xF, false);
    const uint32_t nLA = (uint32_t)__builtin_amdgcn_update_dpp((int)LAu, (int)LAu, kDppWaveShr1, 0xF, 0xF,
                                                               false);
    upM = g > 0 ? nM : kNeg;
    upLA = g > 0 ? nLA : LAu;
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
  // Lane g holds the boundary cell (i, 0) at k0 = kB - t - g (C - 1), so only
  // steps t <= kB have one; its first in-band cell of column L1 (i >= L1 - kB
  // inside its slice) comes at step L1 - kB + G - 1 at the earliest (lane G-1).
  const int top_end = min(KB, tmax);
  const int cap_from = L1min - KB + G - 1;
  const int top_plain_end = min(top_end, cap_from - 1);
  const int mid_end = max(top_end, min(tmax, cap_from - 1));
  int t = 1;
  for (; t <= top_plain_end; ++t) step(t, std::integral_constant<int, kRowTop>{});
  for (; t <= top_end; ++t) step(t, std::integral_constant<int, kRowTopCap>{});
  for (; t <= mid_end; ++t) step(t, std::integral_constant<int, kRowFull>{});
  for (; t <= tmax; ++t) step(t, std::integral_constant<int, kRowCapture>{});
  // Last row: a lane computes rows 1..R only, so when R = L2 its registers
  // still hold row L2 (when R < L2, row L2 has no in-band cell).
  if (has_pair && R == L2) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int j = L2 - KB + base_d + k;
      if (j >= 1 && j <= L1 && base_d + k < W) {
        const int Mr = Mt[k] - 4 * L2;
        if (Mr >= bIM) {
          bIM = Mr;
          bIj = j;
          bILA = LA[k];
        }
      }
    }
  }

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
    if (TR ? !(jM > iM) : iM > jM) {  // see nw_reg_dp
      la = iLA;
      j = ij;
    } else {
      la = jLA;
      j = L1;
    }
    const int AL = (int)(la & 0xFFFFu) + j;
    const int ML = AL - (int)(la >> 16);
    *out = (float)ML / (float)AL;
  }
}

// One wave: 64 / G pairs starting at order[wave * (64 / G)].
template <int C, int KEND>
__device__ __forceinline__ void nw_sliced_wave(const uint32_t *__restrict__ pack,
                                               const int64_t *__restrict__ poff,
                                               const int32_t *__restrict__ len,
                                               const int2 *__restrict__ pairs,
                                               const uint32_t *__restrict__ order, int32_t count,
                                               int KB, int G, int wave, int (*red)[6],
                                               float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int groups = 64 / G;
  const int grp = lane / G;
  const int g = lane - grp * G;
  const int slot = wave * groups + grp;
  const bool has_pair = grp < groups && slot < count;
  const uint32_t pidx = order[has_pair ? slot : count - 1];
  const int2 pr = pairs[pidx];
  const int La = len[pr.x], Lb = len[pr.y];
  const uint32_t *sa = pack + poff[pr.x];
  const uint32_t *sb = pack + poff[pr.y];
  if (wave_transposed(has_pair, La, Lb))
    nw_sliced_dp<C, KEND, true>(sb, sa, Lb, has_pair ? La : 0, has_pair, KB, G, g, red, out + pidx);
  else
    nw_sliced_dp<C, KEND, false>(sa, sb, La, has_pair ? Lb : 0, has_pair, KB, G, g, red, out + pidx);
}

// ------------------------------------------------------------ fused launches
// One launch per kernel family instead of one per bucket: block = one wave =
// one task of the family's list in the device plan.  The tasks run longest
// first, so the grid drains like a longest-processing-time schedule and small
// buckets do not each pay a launch (and a full-wave DP latency) in series.
// The grid is an upper bound on the tasks, set before the histogram exists;
// blocks past the plan's task count exit at once, so the phase needs no host
// read-back.  A block's bucket: the last list entry whose first task is <= its
// task (the entries ascend), found by one ballot per 64 entries.
struct NwTask {
  int kb;     // bucket
  int wave;   // wave index within the bucket
  uint2 bk;   // btab entry of the bucket
};

__device__ __forceinline__ bool nw_task(const NwPlan *__restrict__ plan, const int f, NwTask &task) {
  const uint32_t t = blockIdx.x;
  if (t >= plan->ntasks[f]) return false;
  const int nbk = (int)plan->nbuckets[f];
  const uint2 *list = plan->list[f];
  int k = 0;
  for (int base = 0; base < nbk; base += 64) {
    const int e = base + (int)(threadIdx.x & 63);
    k += __popcll(__ballot(e < nbk && list[e].y <= t));
  }
  const uint2 e = list[k - 1];
  task.kb = (int)e.x;
  task.bk = plan->btab[e.x];
  const uint32_t per = (uint32_t)nw_pairs_per_wave(task.kb, f);
  const uint32_t waves = (task.bk.y + per - 1) / per;
  task.wave = (int)(waves - 1 - (t - e.y));
  return true;
}

template <int KB, int HI>
__device__ __forceinline__ void reg_dispatch(int kb, const uint32_t *pack, const int64_t *poff,
                                             const int32_t *len, const int2 *pairs,
                                             const uint32_t *order, int32_t count, int base,
                                             float *out) {
  if constexpr (KB <= HI) {
    if (kb == KB) {
      nw_reg_wave<KB>(pack, poff, len, pairs, order, count, base, out);
      return;
    }
    reg_dispatch<KB + 1, HI>(kb, pack, poff, len, pairs, order, count, base, out);
  }
}

// Register kernel for buckets LO..HI (family F); WAVES is the smallest
// occupancy target among them (nw_reg_waves).
template <int LO, int HI, int WAVES, int F>
__global__ __launch_bounds__(64, WAVES) void nw_reg_fused_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, const NwPlan *__restrict__ plan, float *__restrict__ out) {
  NwTask task;
  if (!nw_task(plan, F, task)) return;
  reg_dispatch<LO, HI>(task.kb, pack, poff, len, pairs, order + task.bk.x, (int32_t)task.bk.y,
                       task.wave * 64, out);
}

template <int C, int KEND>
__device__ __forceinline__ void sliced_dispatch(int kend, const uint32_t *pack, const int64_t *poff,
                                                const int32_t *len, const int2 *pairs,
                                                const uint32_t *order, int32_t count, int kb, int G,
                                                int wave, int (*red)[6], float *out) {
  if constexpr (KEND < C) {
    if (kend == KEND) {
      nw_sliced_wave<C, KEND>(pack, poff, len, pairs, order, count, kb, G, wave, red, out);
      return;
    }
    sliced_dispatch<C, KEND + 2>(kend, pack, poff, len, pairs, order, count, kb, G, wave, red, out);
  }
}

// Lane-sliced kernel at slice width C (family F) for the buckets assigned to it.
template <int C, int F>
__global__ __launch_bounds__(64) void nw_sliced_fused_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, const NwPlan *__restrict__ plan, float *__restrict__ out) {
  static_assert(kSliceWidths[F - kNumRegFamilies] == C, "family F runs slice width C");
  __shared__ int red[64][6];
  NwTask task;
  if (!nw_task(plan, F, task)) return;
  // KEND = (W - 1) mod C = 2kB mod C, even
  sliced_dispatch<C, 0>((2 * task.kb) % C, pack, poff, len, pairs, order + task.bk.x, (int32_t)task.bk.y,
                        task.kb, slice_lanes(task.kb, C), task.wave, red, out);
}

}  // namespace crispy
