// sort_config.h -- the device radix sort used for the per-slice particle
// re-sort by radial cell and the beam's (xi, r)-cell sort.
//
// rocPRIM's default dispatch sorts up to 2^20 32-bit keys by block sort +
// log2(n / block) merge passes (two launches each: 18 launches for 2^19
// keys).  A merge-sort limit of 0 selects Onesweep instead: one histogram
// pass plus one pass per 8-bit digit, i.e. 2 digit passes for the 13-bit
// cell keys of a 4096-cell slice.  Both are stable, so equal keys keep
// their order and every run is reproducible.
#pragma once

#include <cstring>   // rocPRIM's host code uses memset

#include <rocprim/rocprim.hpp>

namespace qpad {

using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

}  // namespace qpad
