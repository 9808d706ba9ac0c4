// Register-kernel family 1: kB buckets 11..16 in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

static_assert(nw_reg_waves<11>() >= 3 && nw_reg_waves<16>() == 3,
              "family bounds follow nw_reg_waves");

void launch_nw_reg_family1(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_reg_fused_kernel<11, 16, 3, 1><<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order,
      a.plan, a.out);
}

}  // namespace crispy
