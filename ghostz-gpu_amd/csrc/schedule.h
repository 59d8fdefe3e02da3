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

// The learned schedule's two-level keys.  A job's query side length (the
// residues it may extend before the query's delimiter) bounds its DP cells
// and orders most of the work; the cost model orders the jobs whose side is
// long enough for the bound to leave room (at least `gate` residues), inside
// their length bin:
//   key = min(length, 63) << 6 | prediction bin,
// the prediction bin being clamp(floor(4 * predicted log2 cells), 0, 63) for
// a gated job and 0 otherwise.
constexpr uint32_t kLengthBins = 64;
// keys[j] = min(length[j], 63) << 6 for j in [0, n), and the indices j with
// length[j] >= gate appended to index (count at *d_count, which the call
// clears first); the order of the list is unspecified.
hipError_t BinAndGateJobs(const uint16_t *length, uint32_t n, uint32_t gate, uint16_t *keys,
                          uint32_t *index, uint32_t *d_count, hipStream_t stream);
// keys[index[k]] |= prediction bin of pred[k], k < m (k < *d_count when
// d_count is given).
hipError_t AddPredictionBins(const float *pred, const uint32_t *index, uint32_t m,
                             uint16_t *keys, hipStream_t stream,
                             const uint32_t *d_count = nullptr);

}  // namespace gz
