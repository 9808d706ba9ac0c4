// Register-kernel family 3: kB buckets 25..32 in one fused launch.
#include "nw_dp.hpp"

namespace crispy {

static_assert(nw_reg_waves<25>() >= 2 && nw_reg_waves<kKbRegMax>() == 2,
              "family bounds follow nw_reg_waves");

void launch_nw_reg_family3(hipStream_t s, unsigned blocks, const NwFusedArgs &a) {
  nw_reg_fused_kernel<25, kKbRegMax, 2, 3><<<blocks, 64, 0, s>>>(a.pack, a.poff, a.len, a.pairs, a.order,
      a.plan, a.out);
}

}  // namespace crispy
