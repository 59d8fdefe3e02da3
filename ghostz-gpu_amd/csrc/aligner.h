// aligner.h -- AlignerGpu for one device (src/aligner_gpu.{h,cpp},
// src/aligner_gpu_presearch_thread.cpp): the database chunks stay resident
// in HBM; each query chunk runs
//   seed stage (device) -> per-query order + chain filter (host)
//   -> gapped X-drop both sides (device) -> AddResults (host)
// per database chunk, then BuildResults (host traceback) and WriteOutput.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "database.h"
#include "queries.h"
#include "results.h"
#include "search_device.h"

struct gz_gapped_extender_s;

namespace gz {

struct AlignStats {
  double seconds_total = 0;       // presearch start .. output written
  double seconds_query_lists = 0;
  double seconds_upload = 0;
  double seconds_seed = 0;        // device seed stage incl. hit download
  double seconds_chain = 0;
  double seconds_gapped = 0;      // device gapped extension incl. transfers
  double seconds_add = 0;
  double seconds_build = 0;
  double seconds_output = 0;
  uint64_t query_chunks = 0;
  uint64_t lists = 0;
  uint64_t units_rep = 0, units_value = 0;
  uint64_t seed_hits = 0;
  uint64_t chained_hits = 0;
  uint64_t gapped_extensions = 0;
  uint64_t gapped_cells = 0;
  uint64_t results = 0;
  uint64_t sequential_build_chunks = 0;
  double gapped_kernel_ms = 0;    // measured with HIP events on the stream
};

enum class Schedule { kPlain = 0, kLengthSorted = 1, kLearned = 2 };

class CostModel;  // learned gapped-extension cost (cost_model.h)

struct AlignerOptions {
  SearchParameters search;
  uint64_t queries_chunk_size = 1ull << 27;  // -l (framed codes)
  int threads = 0;
  Schedule schedule = Schedule::kPlain;
  const CostModel *cost_model = nullptr;     // for Schedule::kLearned
  bool record_cells = false;                 // keep per-extension DP cells
};

// One gapped extension job as traced for the learned scheduler.
struct GappedTrace {
  uint32_t qpos;   // concatenated seed position
  uint32_t dpos;
  uint32_t chunk;
  uint8_t reverse;
  uint32_t cells;
};

class Aligner {
 public:
  explicit Aligner(const AlignerOptions &opt);
  ~Aligner();
  hipError_t SetDatabase(const Database *db);
  // Aligns all queries (split into chunks by queries_chunk_size) and appends
  // the tabular output.
  hipError_t Align(const QuerySet &queries, std::string *output, AlignStats *stats);
  const std::vector<GappedTrace> &Traces() const { return traces_; }
  const std::vector<uint8_t> &LastConcatenatedQueries() const { return concat_; }

 private:
  struct DeviceChunk;
  hipError_t AlignChunk(const QuerySet &q, std::string *output, AlignStats *st);
  AlignerOptions opt_;
  const Database *db_ = nullptr;
  std::vector<std::unique_ptr<DeviceChunk> > chunks_;
  hipStream_t stream_ = nullptr;
  int *d_matrix_ = nullptr;
  uint8_t *d_reduced_ = nullptr;
  SeedStage seed_;
  gz_gapped_extender_s *gapped_ = nullptr;
  std::vector<GappedTrace> traces_;
  std::vector<uint8_t> concat_;
};

}  // namespace gz
