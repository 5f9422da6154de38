// sort.hpp -- integer scan and particle bookkeeping shared by the drop-in
// entry points and resident simulations.
#pragma once

#include "common.hpp"
#include "pic_device.hpp"

namespace chimera {

// Exclusive prefix sum of n int32 values into int64 out; the grand total is
// written to *total_dev (device memory, may be null).  Uses scratch slot
// scratch_slot for the tile sums.
int32_t exclusive_scan(const int32_t* in, int64_t* out, int64_t n, int64_t* total_dev, int scratch_slot);

// Stable LSD radix sort of n keys of key_bits bits; idx receives the source
// index of every sorted key.  keys_tmp/idx_tmp are n-element work buffers.
int32_t radix_sort_keys(uint32_t* keys, int32_t* idx, uint32_t* keys_tmp, int32_t* idx_tmp, int64_t n, int key_bits,
                        int scratch_hist, int scratch_scan);

}  // namespace chimera
