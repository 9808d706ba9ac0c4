// Register-kernel family 0: kB buckets 0..10 in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

static_assert(nw_reg_waves<0>() >= 4 && nw_reg_waves<10>() == 4,
              "family bounds follow nw_reg_waves");

void launch_nw_reg_family0(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_reg_fused_kernel<0, 10, 4, 0><<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order,
      a.plan, a.out);
}

}  // namespace crispy
