// Fused NW DP launches (one per kernel family), defined in nw_fam_*.hip.
// Block = one wave = one task (bucket kB, wave index in the bucket) of the
// family's list in the device plan; the grid is an upper bound on the tasks,
// so it does not depend on the bucket histogram.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace crispy {

// Register-kernel families: bucket ranges sharing one occupancy target
// (3 VGPRs per band cell + ~64: 4 waves/SIMD up to kB 10, 3 up to 16, 2 above);
// kB 17..32 is split in two only to keep each translation unit's compile short.
struct RegFamily {
  int lo, hi;
};
constexpr RegFamily kRegFamilies[] = {{0, 10}, {11, 16}, {17, 24}, {25, kKbRegMax}};
constexpr int kNumRegFamilies = 4;
static_assert(kRegFamilies[0].lo == 0 && kRegFamilies[0].hi + 1 == kRegFamilies[1].lo &&
                  kRegFamilies[1].hi + 1 == kRegFamilies[2].lo &&
                  kRegFamilies[2].hi + 1 == kRegFamilies[3].lo && kRegFamilies[3].hi == kKbRegMax,
              "the register families cover kB 0..kKbRegMax without gaps");
// Lane-sliced families: G = ceil(W / C) lanes per pair, C band cells per lane,
// one family per slice width.  Which width a bucket runs at comes from
// slice_width_for().
constexpr int kNumSliceWidths = 2;
constexpr int kSliceWidths[kNumSliceWidths] = {24, 32};
constexpr int kNumFamilies = kNumRegFamilies + kNumSliceWidths;

__host__ __device__ constexpr int slice_lanes(int kb, int C) { return (2 * kb + 1 + C - 1) / C; }

// Cost model of one wave step of the sliced kernel: ~13.5 C VALU ops for the C
// cells of every lane plus ~45 per step (match words, neighbour exchange,
// window slide), shared by the 64 / G pairs of the wave.  Measured on MI355X
// the model holds for wide bands but not for G <= 6 lanes at 24 cells (kB up
// to ~71: 24 cells ran 3 % faster on kB 33..55 where the model favours 32), so
// the wider slice is taken only from 7 lanes up.  `forced` (24 or 32)
// overrides the choice.  Integer form (cost x 2, cross-multiplied), so the
// host and the device plan agree exactly.
__host__ __device__ inline int slice_width_for(int kb, int forced) {
  if (forced == 24 || forced == 32) return forced;
  const int g24 = slice_lanes(kb, 24), p24 = 64 / g24, p32 = 64 / slice_lanes(kb, 32);
  // c32 <= c24  <=>  (13.5*32 + 45) / p32 <= (13.5*24 + 45) / p24
  return g24 >= 7 && (27 * 32 + 90) * p24 <= (27 * 24 + 90) * p32 ? 32 : 24;
}

// Kernel family of bucket kB (kB <= kKbSliceMax) under slice-width policy
// `forced` (0 = cost model).
__host__ __device__ inline int nw_family_of(int kb, int forced) {
  if (kb > kKbRegMax) return kNumRegFamilies + (slice_width_for(kb, forced) == kSliceWidths[0] ? 0 : 1);
  int f = 0;
  while (kb > kRegFamilies[f].hi) ++f;
  return f;
}

// Pairs one wave task of bucket kB covers in family f.
__host__ __device__ inline int nw_pairs_per_wave(int kb, int f) {
  return f >= kNumRegFamilies ? 64 / slice_lanes(kb, kSliceWidths[f - kNumRegFamilies]) : 64;
}

// Device plan of one DP phase, written by nw_plan_kernel from the bucket
// histogram.  btab[b] = (first position in `order`, pairs) of bucket b.  Per
// family f: its non-empty buckets in task order, list[f][k] = (bucket, first
// task), ntasks[f] tasks in all; the tasks of a family run its buckets by
// descending mean cells per pair and a bucket's waves from its longest pairs
// (the end of the R order) down, so the grid drains like a
// longest-processing-time schedule.
struct NwPlan {
  uint2 btab[kNumBuckets];
  uint2 list[kNumFamilies][kNumBuckets];
  uint32_t ntasks[kNumFamilies];
  uint32_t nbuckets[kNumFamilies];
};

struct NwFusedArgs {
  const uint32_t *pack;
  const int64_t *poff;
  const int32_t *len;
  const int2 *pairs;
  const uint32_t *order;
  const NwPlan *plan;
  float *out;
};

// Each launcher runs `blocks` one-wave blocks of its family.
void launch_nw_reg_family0(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family1(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family2(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family3(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_sliced_family24(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_sliced_family32(hipStream_t s, unsigned blocks, const NwFusedArgs &a);

}  // namespace crispy
