// reference_shims.h -- the class-level drop-in a maintainer adds to the
// reference in place of its CUDA classes: the reference's class and method
// names (src/distance_calculator_gpu.h:18-40, src/ungapped_extender_gpu.h
// :21-76, src/gapped_extender_gpu.h:19-60, src/host_seeds_memory.h:18-80,
// src/gpu_stream_controller.h:24-101, src/seed_searcher_gpu.h:28-115) over
// the C ABI of include/ghostz_hip.h.  `cudaStream_t &` becomes
// `gz_stream_t &`; AlphabetCoder::Code is uint8_t and PackedAlphabetCode is
// uint64_t in the reference, so device buffers pass through unchanged.  As in
// the reference, a device error ends the process (src/cuda_common.h:13-19).
//
// Compiled with g++ against the header alone and linked with
// libghostz_amd.so by tests/test_reference_shims.py.
#pragma once

#include <stdint.h>
#include <stdlib.h>

#include <iostream>
#include <vector>

#include "ghostz_hip.h"

namespace shim {

inline int Check(gz_status s) {
  if (s != GZ_OK) {
    std::cerr << "ghostz_hip: " << gz_last_error() << std::endl;
    exit(1);
  }
  return 0;
}

class DistanceCalculatorGpu {
 public:
  typedef char Distance;
  DistanceCalculatorGpu() { Check(gz_distance_calculator_create(&h_)); }
  ~DistanceCalculatorGpu() { gz_distance_calculator_destroy(h_); }
  int SetQueries(const uint64_t *d_queries) {
    return Check(gz_distance_calculator_set_queries(h_, d_queries));
  }
  int SetDatabase(const uint64_t *d_database) {
    return Check(gz_distance_calculator_set_database(h_, d_database));
  }
  int SetReducedCodeMap(const uint8_t *d_map) {
    return Check(gz_distance_calculator_set_reduced_code_map(h_, d_map));
  }
  int SetSubsequenceLength(uint32_t length) {
    return Check(gz_distance_calculator_set_subsequence_length(h_, length));
  }
  int CalculateDistancesAsync(size_t size, uint32_t *h_qcenters, uint32_t *h_dbcenters,
                              Distance *h_dist, uint32_t *d_qcenters, uint32_t *d_dbcenters,
                              Distance *d_dist, gz_stream_t &stream) {
    return Check(gz_distance_calculator_calculate_distances_async(
        h_, size, h_qcenters, h_dbcenters, h_dist, d_qcenters, d_dbcenters, d_dist, stream));
  }
  gz_distance_calculator_t handle() { return h_; }

 private:
  gz_distance_calculator_t h_ = nullptr;
};

class UngappedExtenderGpu {
 public:
  UngappedExtenderGpu() { Check(gz_ungapped_extender_create(&h_)); }
  ~UngappedExtenderGpu() { gz_ungapped_extender_destroy(h_); }
  int SetQuerySeedDataList(const uint32_t *d_query_ids, const uint32_t *d_query_positions) {
    return Check(gz_ungapped_extender_set_query_seed_data_list(h_, d_query_ids,
                                                               d_query_positions));
  }
  int SetQueries(uint8_t delimiter, const uint64_t *d_queries, const int *d_cutoffs,
                 const int *d_triggers) {
    return Check(gz_ungapped_extender_set_queries(h_, delimiter, d_queries, d_cutoffs,
                                                  d_triggers));
  }
  int SetDatabase(const uint64_t *d_database) {
    return Check(gz_ungapped_extender_set_database(h_, d_database));
  }
  int SetScoreMatrix(const int *d_matrix, uint32_t number_letters) {
    return Check(gz_ungapped_extender_set_score_matrix(h_, d_matrix, number_letters));
  }
  gz_ungapped_extender_t handle() { return h_; }

 private:
  gz_ungapped_extender_t h_ = nullptr;
};

class GappedExtenderGpu {
 public:
  GappedExtenderGpu() { Check(gz_gapped_extender_create(&h_)); }
  ~GappedExtenderGpu() { gz_gapped_extender_destroy(h_); }
  int SetQueries(uint8_t delimiter, const uint64_t *d_queries) {
    return Check(gz_gapped_extender_set_queries(h_, delimiter, d_queries));
  }
  int SetDatabase(const uint64_t *d_database) {
    return Check(gz_gapped_extender_set_database(h_, d_database));
  }
  int SetScoreParameters(const int *d_matrix, uint32_t number_letters, int gap_open,
                         int gap_extension, int cutoff) {
    return Check(gz_gapped_extender_set_score_parameters(h_, d_matrix, number_letters,
                                                         gap_open, gap_extension, cutoff));
  }
  gz_gapped_extender_t handle() { return h_; }

 private:
  gz_gapped_extender_t h_ = nullptr;
};

class HostSeedsMemory {
 public:
  HostSeedsMemory(size_t id, size_t size) : id_(id) {
    Check(gz_host_seeds_memory_create(size, &h_));
  }
  ~HostSeedsMemory() { gz_host_seeds_memory_destroy(h_); }
  size_t GetId() const { return id_; }
  size_t GetSize() const { return gz_host_seeds_memory_size(h_); }
  uint32_t *GetQuerySeedIds() { return gz_host_seeds_memory_query_seed_ids(h_); }
  uint32_t *GetQuerySeedStarts() { return gz_host_seeds_memory_query_seed_starts(h_); }
  uint32_t *GetQueryIds() { return gz_host_seeds_memory_query_ids(h_); }
  uint32_t *GetQueryConcatenatedPositions() {
    return gz_host_seeds_memory_query_concatenated_positions(h_);
  }
  uint32_t *GetDatabasePositions() { return gz_host_seeds_memory_database_positions(h_); }
  int *GetResultValues() { return gz_host_seeds_memory_result_values(h_); }
  gz_host_seeds_memory_t handle() { return h_; }

 private:
  size_t id_;
  gz_host_seeds_memory_t h_ = nullptr;
};

class GpuStreamController {
 public:
  static const size_t kMaxNumberOfConcurrentExecutions = 2;
  struct DistanceCalculationSeedsParameters {
    size_t size;
  };
  struct UngappedExtensionWithTriggerSeedsParameters {
    size_t query_seed_size;
    size_t size;
  };
  struct GappedExtensionSeesParameters {
    size_t reverse_offset;
    size_t reverse_length;
    size_t forward_offset;
    size_t forward_length;
  };
  GpuStreamController() { Check(gz_stream_controller_create(&h_)); }
  ~GpuStreamController() { gz_stream_controller_destroy(h_); }
  void SetDistanceCalculator(DistanceCalculatorGpu *d) {
    Check(gz_stream_controller_set_distance_calculator(h_, d->handle()));
  }
  void SetUngappedExtender(UngappedExtenderGpu *u) {
    Check(gz_stream_controller_set_ungapped_extender(h_, u->handle()));
  }
  void SetGappedExtender(GappedExtenderGpu *g) {
    Check(gz_stream_controller_set_gapped_extender(h_, g->handle()));
  }
  int Init(size_t device_buffer_size) {
    return Check(gz_stream_controller_init(h_, device_buffer_size));
  }
  bool FinishedRun(int resource_id) {
    return gz_stream_controller_finished_run(h_, resource_id) != 0;
  }
  void WaitRun() { Check(gz_stream_controller_wait_run_all(h_)); }
  void WaitRun(int resource_id) { Check(gz_stream_controller_wait_run(h_, resource_id)); }
  int Run(DistanceCalculationSeedsParameters &p, HostSeedsMemory &m) {
    int id = -1;
    Check(gz_stream_controller_run_distance_calculation(h_, p.size, m.handle(), &id));
    return id;
  }
  int Run(UngappedExtensionWithTriggerSeedsParameters &p, HostSeedsMemory &m) {
    int id = -1;
    Check(gz_stream_controller_run_ungapped_extension_with_trigger(h_, p.query_seed_size,
                                                                   p.size, m.handle(), &id));
    return id;
  }
  int Run(GappedExtensionSeesParameters &p, HostSeedsMemory &m) {
    int id = -1;
    Check(gz_stream_controller_run_gapped_extension(h_, p.reverse_offset, p.reverse_length,
                                                    p.forward_offset, p.forward_length,
                                                    m.handle(), &id));
    return id;
  }
  void ReleaseResource(int resource_id) {
    Check(gz_stream_controller_release_resource(h_, resource_id));
  }
  uint32_t GetNumberOfAvailableResources() {
    return gz_stream_controller_available_resources(h_);
  }

 private:
  gz_stream_controller_t h_ = nullptr;
};

// SeedSearcherGpu: Reset takes the query set and one database chunk (the
// reference passes the query / database parameter objects built from them);
// the searches return the passing seeds of a range of queries.
class SeedSearcherGpu {
 public:
  struct Hit {
    uint32_t query_sequence_position;
    uint32_t database_sequence_position;
  };
  explicit SeedSearcherGpu(const gz_aligner_parameters *p = nullptr) {
    Check(gz_seed_searcher_create(p, &h_));
  }
  ~SeedSearcherGpu() { gz_seed_searcher_destroy(h_); }
  uint32_t Reset(gz_queries_t queries, gz_database_t database, uint32_t chunk) {
    uint32_t lists = 0;
    Check(gz_seed_searcher_reset(h_, queries, database, chunk, &lists));
    number_lists_ = lists;
    return 0;
  }
  uint32_t GetNumberHashPositionDataLists() const { return number_lists_; }
  // the distance-calculation pairs of one list (query concatenated
  // position, representative position)
  int SearchSeedsFromRepresentativesForSimilarityFiltering(
      uint32_t list, std::vector<uint32_t> *query_positions,
      std::vector<uint32_t> *database_positions) const {
    size_t n = 0;
    gz_status s = gz_seed_searcher_search_seeds_from_representatives_for_similarity_filtering(
        h_, list, nullptr, nullptr, 0, &n);
    if (s != GZ_OK && s != GZ_ERROR_CAPACITY) Check(s);
    query_positions->resize(n);
    database_positions->resize(n);
    if (n)
      Check(gz_seed_searcher_search_seeds_from_representatives_for_similarity_filtering(
          h_, list, query_positions->data(), database_positions->data(), n, &n));
    return 0;
  }
  int SearchSeedsFromClusteringSubsequence(uint32_t qb, uint32_t qe,
                                           std::vector<std::vector<Hit> > *hits) const {
    gz_seed_hits out;
    Check(gz_seed_searcher_search_seeds_from_clustering_subsequence(h_, qb, qe, &out));
    Append(out, qb, hits);
    return 0;
  }
  int SearchSeedsFromNonClusteringSubsequence(uint32_t qb, uint32_t qe,
                                              std::vector<std::vector<Hit> > *hits) const {
    gz_seed_hits out;
    Check(gz_seed_searcher_search_seeds_from_non_clustering_subsequence(h_, qb, qe, &out));
    Append(out, qb, hits);
    return 0;
  }
  int SearchSeeds(uint32_t qb, uint32_t qe, std::vector<std::vector<Hit> > *hits) const {
    gz_seed_hits out;
    Check(gz_seed_searcher_search_seeds(h_, qb, qe, &out));
    Append(out, qb, hits);
    return 0;
  }

 private:
  static void Append(const gz_seed_hits &out, uint32_t qb,
                     std::vector<std::vector<Hit> > *hits) {
    if (hits->size() < qb + out.number_queries) hits->resize(qb + out.number_queries);
    for (uint32_t i = 0; i < out.number_queries; ++i)
      for (uint32_t k = out.query_offsets[i]; k < out.query_offsets[i + 1]; ++k)
        (*hits)[qb + i].push_back({out.query_positions[k], out.database_positions[k]});
  }
  gz_seed_searcher_t h_ = nullptr;
  uint32_t number_lists_ = 0;
};

}  // namespace shim
