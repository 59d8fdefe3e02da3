// traceback.h -- the result build's traceback re-extensions on the device.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include "presearch.h"

namespace gz {

// Per-lane scratch capacities (elements) and the lane stride in bytes.
struct LaneCaps {
  uint32_t dp = 0, rows = 0, arena = 0, ops = 0;
  size_t bytes_per_lane = 0;
};

// The op arena of a lane.
constexpr uint32_t kTracebackArena = 1u << 16;

LaneCaps TracebackCaps(uint32_t max_query_length, uint32_t max_subject_length,
                       uint32_t arena = kTracebackArena);

// One result per job; `lanes` (a multiple of 64) is the number of lanes the
// scratch holds (lanes x caps.bytes_per_lane bytes).
hipError_t TracebackLaunch(const uint64_t *packed_queries, const uint64_t *packed_database,
                           const TracebackJob *jobs, uint32_t n, const int *matrix,
                           uint32_t nl, int gap_open, int gap_ext, int cutoff,
                           uint8_t *scratch, const LaneCaps &caps, uint32_t lanes,
                           TracebackOut *outs, hipStream_t stream);

}  // namespace gz
