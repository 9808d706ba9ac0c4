// Lane-sliced family at 32 band cells per lane: the kB buckets above
// kKbRegMax that slice_width_for() assigns to this width, in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

void launch_nw_sliced_family32(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_sliced_fused_kernel<32, kNumRegFamilies + 1><<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order,
      a.plan, a.out);
}

}  // namespace crispy
