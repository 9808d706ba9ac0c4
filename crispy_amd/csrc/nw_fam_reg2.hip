// Register-kernel family 2: kB buckets 17..24 in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

static_assert(nw_reg_waves<17>() >= 2 && nw_reg_waves<24>() == 2,
              "family bounds follow nw_reg_waves");

void launch_nw_reg_family2(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_reg_fused_kernel<17, 24, 2, 2><<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order,
      a.plan, a.out);
}

}  // namespace crispy
