// Lane-sliced family: kB buckets kKbRegMax+1 .. kKbSliceMax in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

__global__ __launch_bounds__(64) void nw_sliced_fused_kernel(
    const uint32_t *__restrict__ pack, const int64_t *__restrict__ poff,
    const int32_t *__restrict__ len, const int2 *__restrict__ pairs,
    const uint32_t *__restrict__ order, const uint2 *__restrict__ btab,
    const uint2 *__restrict__ tasks, float *__restrict__ out) {
  __shared__ int red[64][6];
  const uint2 task = tasks[blockIdx.x];
  const uint2 bk = btab[task.x];
  const int kb = (int)task.x;
  const int G = (2 * kb + 1 + kSliceCells - 1) / kSliceCells;
  // KEND = (W - 1) mod 24 = 2kB mod 24, even
  sliced_dispatch<0>((2 * kb) % kSliceCells, pack, poff, len, pairs, order + bk.x, (int32_t)bk.y,
                     kb, G, (int)task.y, red, out);
}


void launch_nw_sliced_family(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_sliced_fused_kernel<<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order, a.btab,
                                               a.tasks, a.out);
}

}  // namespace crispy
