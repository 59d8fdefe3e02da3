// radix_sort.h -- stable device sort of (u32 key, u32 value) pairs, the
// project's own replacement for a library radix sort.
//
// LSD passes of 8 bits.  A pass: per-tile digit histograms (tiles of 4096
// items, digit-major so that one exclusive scan gives every (digit, tile)
// its output offset), then a scatter in which each wave ranks its 64 items
// among equal digits with eight ballots, and the waves' counts are
// prefix-summed in LDS in item order, so equal keys keep their input order.
#pragma once

#include <stdint.h>
#include <stddef.h>

#include <hip/hip_runtime.h>

namespace gz {

// Device scratch bytes for sorting n pairs (besides the alternate buffers).
size_t RadixSortTempBytes(size_t n);

// Sorts keys[0, n) ascending over bits [0, bits) (bits 16 or 32: an even
// number of passes), carrying values along, stably.  keys_alt / values_alt are
// scratch of n items each; the result ends in keys / values.
hipError_t RadixSortPairs(uint32_t *keys, uint32_t *values, uint32_t *keys_alt,
                          uint32_t *values_alt, size_t n, int bits, void *temp,
                          hipStream_t stream);

}  // namespace gz
