// schedule.h -- device-side ordering of independent extension jobs.
//
// The gapped extensions of a chunk are independent: any order gives the same
// results (each job's outputs are scattered back to its own slot).  Ordering
// only changes how much work shares a wavefront.  Jobs are binned by a
// 16-bit key (larger = more predicted work) and laid out by descending bin
// with a three-kernel counting sort (block histograms, bin scan, block-ranked
// scatter); the order inside a bin is unspecified.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace gz {

constexpr uint32_t kScheduleBins = 4096;

// Device workspace bytes for ordering n jobs.
size_t ScheduleTempBytes(size_t n);

// perm[k] = index of the job placed k-th (descending key; keys < kScheduleBins).
hipError_t DeviceOrderByKey(const uint16_t *keys, uint32_t n, uint32_t *perm, void *temp,
                            hipStream_t stream);

// key = clamp(round(x * scale), 0, kScheduleBins - 1) of a predicted cost.
hipError_t CostToKey(const float *cost, uint32_t n, float scale, uint16_t *keys,
                     hipStream_t stream);

}  // namespace gz
