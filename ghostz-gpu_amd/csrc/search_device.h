// search_device.h -- the device-resident seed stage of the presearch.
//
// Replaces the host enumeration of SeedSearcherGpu
// (src/seed_searcher_gpu.cpp:129-534), the DistanceCalculationSeedList /
// UngappedExtensionWithTriggerSeedList batches and the host flag compaction
// (src/aligner_gpu_presearch_thread.cpp:822-896) with device kernels:
//   lookup:   query hash list -> database representatives / values
//   units:    (list, representative, entry) and (list, value, entry) work
//             units, prefix-summed and mapped to lanes by load-balanced search
//   seeds:    per unit: reduced-alphabet distance, ungapped X-drop with
//             trigger for the representative, its members, or the value
//   append:   wave-aggregated (ballot/popcount) append of passing seeds
// Each passing seed carries the key (list rank, representative or value
// index, member index + 1 or 0, entry index) that orders it exactly as the
// one-thread CPU search enumerates it.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <mutex>

#include <vector>

#include "presearch.h"

namespace gz {

struct DeviceDbChunk {
  const uint64_t *packed = nullptr;
  uint32_t nhash = 0;
  const uint32_t *hashes = nullptr, *rep_off = nullptr, *rep_pos = nullptr,
                 *mem_off = nullptr, *mem_pos = nullptr, *val_off = nullptr,
                 *val_pos = nullptr;
};

struct DeviceQueryLists {
  const uint64_t *packed = nullptr;  // concatenated queries
  const int *cutoffs = nullptr;
  const int *triggers = nullptr;
  uint32_t nqueries = 0;
  uint32_t nlists = 0;
  const uint32_t *hash = nullptr, *entry_off = nullptr, *entry_pos = nullptr,
                 *entry_qid = nullptr;
  const uint8_t *entry_check = nullptr;
  const uint4 *entry_rec = nullptr;  // per entry: position, query | check << 31, cutoff, trigger
};

struct SeedStageParams {
  const int *matrix = nullptr;  // device nl x nl
  uint32_t nl = 25;
  uint32_t delimiter = 25;
  const uint8_t *reduced_map = nullptr;  // device, 25 entries
  uint32_t subsequence_length = 10;
  int distance_threshold = 2;  // rep extension when d <= threshold
  int max_mismatches = 1;      // members when max(d - mm, 0) <= threshold
  bool clustering = true;
  // byte ungapped score tables: every cutoff of the query chunk in
  // [1, kMaxByteCutoff] and every matrix score in [-127, 127]; 16-bit tables
  // otherwise
  bool byte_tables = false;
};

struct SeedStageStats {
  uint64_t units_rep = 0;      // (list, rep, entry) units
  uint64_t units_value = 0;    // (list, value, entry) units
  uint64_t distance_pairs = 0;
  uint64_t hits = 0;
  // seed kernel time (HIP events on the stream) and launches, per kind
  double kernel_ms[2] = {0, 0};  // [0] representatives, [1] values
  uint64_t launches[2] = {0, 0};
};

// Workspace owned by the caller (grown on demand).
class SeedStage {
 public:
  ~SeedStage();
  // Builds the per-entry records the seed kernels read (one 16-byte load
  // per unit) and sets q->entry_rec; once per query chunk.
  hipError_t PrepareEntries(DeviceQueryLists *q, uint64_t entries, hipStream_t stream);
  // Finds the database representatives and values of every list of a query
  // chunk in one database chunk; Run uses them until the next Lookup.
  hipError_t Lookup(const DeviceQueryLists &q, const DeviceDbChunk &db, bool clustering,
                    hipStream_t stream);
  // Runs the seed stage of queries [qb, qe) of the chunk (qb a multiple of
  // 64, qe too or the number of queries) against the looked-up database
  // chunk and returns their passing seeds grouped by query (qstart indexed
  // by qid - qb), in pinned host memory owned by the stage.  The stage keeps
  // two such buffers and alternates: a view stays valid until the second
  // Run after it.  kinds selects the seed sources: bit 0 the representatives
  // and their members, bit 1 the non-clustered values.
  hipError_t Run(const DeviceQueryLists &q, const DeviceDbChunk &db,
                 const SeedStageParams &p, uint32_t qb, uint32_t qe, hipStream_t stream,
                 SeedHitView *hits, SeedStageStats *stats, unsigned kinds = 3,
                 bool compact = false);

 private:
  hipError_t Reserve(void **p, size_t *cap, size_t bytes);
  void *lookup_ = nullptr; size_t lookup_cap_ = 0;
  void *units_ = nullptr; size_t units_cap_ = 0;
  void *scan_tmp_ = nullptr; size_t scan_tmp_cap_ = 0;
  void *active_ = nullptr; size_t active_cap_ = 0;
  void *hits_ = nullptr; size_t hits_cap_ = 0;
  // hits grouped by query, per host buffer (a compact view's records stay
  // on the device until that buffer's next Run)
  void *grouped_[2] = {nullptr, nullptr}; size_t grouped_cap_[2] = {0, 0};
  void *positions_ = nullptr; size_t positions_cap_ = 0;  // compact views' positions
  hipStream_t fetch_stream_ = nullptr;  // copies of a compact view's records
  void *fetch_host_ = nullptr; size_t fetch_host_cap_ = 0;  // their pinned staging
  std::mutex fetch_mutex_;
  void *tile_range_ = nullptr; size_t tile_range_cap_ = 0;  // per-tile list ranges
  void *qcount_ = nullptr; size_t qcount_cap_ = 0;     // per-query counts / cursors
  void *base_info_ = nullptr; size_t base_info_cap_ = 0;  // per-list database ranges
  void *entry_rec_ = nullptr; size_t entry_rec_cap_ = 0;  // per-entry records
  bool looked_up_ = false;
  int flip_ = 0;  // host buffer of the next Run
  void *host_hits_[2] = {nullptr, nullptr};  // pinned
  size_t host_hits_cap_[2] = {0, 0};
  void *host_qstart_[2] = {nullptr, nullptr};  // pinned
  size_t host_qstart_cap_[2] = {0, 0};
  uint32_t *counters_ = nullptr;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> launch_ev_;  // a pair per seed-kernel launch of a run
  void *pinned_ = nullptr; size_t pinned_cap_ = 0;
};

}  // namespace gz
