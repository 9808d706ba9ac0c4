// Radix sort configuration shared by the library's rocPRIM sorts.
#pragma once
#include <rocprim/device/device_radix_sort.hpp>

namespace crispy {

// Onesweep (one histogram pass + one scatter pass per 8-bit digit) from 64 Ki
// items up: for the 10^5 - 10^7 keys of a step rocPRIM's default picks block
// sort + log2(n / block) merge passes, about 20 dispatches.
using RadixOnesweep =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 1 << 16>;

}  // namespace crispy
