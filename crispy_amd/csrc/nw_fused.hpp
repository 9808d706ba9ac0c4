// Fused NW DP launches (one per kernel family), defined in nw_fam_*.hip.
// Block = one wave; tasks[blockIdx.x] = (bucket kB, wave index in the bucket);
// btab[kB] = (first position in `order`, pairs in the bucket).
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
constexpr int kNumFamilies = kNumRegFamilies + 1;  // + the sliced family
constexpr int kSliceCells = 24;                    // band cells per lane, sliced kernel

struct NwFusedArgs {
  const uint32_t *pack;
  const int64_t *poff;
  const int32_t *len;
  const int2 *pairs;
  const uint32_t *order;
  const uint2 *btab;
  const uint2 *tasks;
  float *out;
};

void launch_nw_reg_family0(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family1(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family2(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_reg_family3(hipStream_t s, unsigned blocks, const NwFusedArgs &a);
void launch_nw_sliced_family(hipStream_t s, unsigned blocks, const NwFusedArgs &a);

}  // namespace crispy
