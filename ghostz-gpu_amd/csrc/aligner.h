// aligner.h -- AlignerGpu for one device (src/aligner_gpu.{h,cpp},
// src/aligner_gpu_presearch_thread.cpp): the database chunks stay resident
// in HBM; each query chunk runs
//   seed stage (device) -> per-query order + chain filter (host)
//   -> gapped X-drop both sides (device) -> AddResults (host)
// per database chunk, then BuildResults (host traceback) and WriteOutput.
// The host flow is AlignQueries (presearch.h); this class supplies its
// device stages.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "database.h"
#include "presearch.h"
#include "queries.h"
#include "results.h"
#include "query_lists_device.h"
#include "search_device.h"

struct gz_gapped_extender_s;

namespace gz {

// The batched form's autotuner: per query-length bucket, every
// configuration of the first gapped ring (width 32 or 16 columns, persistent
// work queue or static grid) is timed kTrials times on real blocks, then the
// one with the fewest kernel milliseconds per job is used.  Results never
// depend on the configuration.
struct GappedTuner {
  static constexpr int kConfigs = 4;
  static constexpr int kTrials = 2;
  double ms[kLengthBuckets][kConfigs] = {};
  double jobs[kLengthBuckets][kConfigs] = {};
  int trials[kLengthBuckets][kConfigs] = {};
  static uint32_t Narrow(int config) { return (config & 1) ? 16u : 32u; }
  static int Persistent(int config) { return (config & 2) ? 0 : 1; }
  int Pick(int bucket) const;
  void Record(int bucket, int config, double kernel_ms, double njobs);
  // The task's layout: 0 = every bucket launched on its own with its
  // configuration, 1 = one launch per side over all of the task's jobs (the
  // fewest launches, each waiting once for its longest job).  Layout 0 is
  // timed only on tasks whose buckets all ran their chosen configuration,
  // layout 1 after that; then the faster per job is kept.
  double layout_ms[2] = {};
  double layout_jobs[2] = {};
  int layout_trials[2] = {};
  int PickLayout() const;
  void RecordLayout(int layout, double kernel_ms, double njobs);
  // the configuration a whole-side launch runs (32-column ring, static grid)
  static constexpr int kWholeConfig = 2;
};

class Aligner : public PresearchDevice {
 public:
  explicit Aligner(const AlignerOptions &opt);
  ~Aligner() override;
  hipError_t SetDatabase(const Database *db);
  bool HasDatabase() const { return db_ != nullptr; }
  // Cost model for Schedule::kLearned (not owned; nullptr = none).
  void SetCostModel(const CostModel *model) { opt_.cost_model = model; }
  // Aligns all queries (split into chunks by queries_chunk_size) and appends
  // the tabular output.
  hipError_t Align(const QuerySet &queries, std::string *output, AlignStats *stats);
  const std::vector<GappedTrace> &Traces() const { return traces_; }
  // The resident query chunk's hash lists on the device (after
  // BeginQueryChunk) and the seed sources SeedHits runs (bit 0:
  // representatives and members, bit 1: non-clustered values; default both).
  const DeviceQueryLists &QueryLists() const { return dq_; }
  void SetSeedKinds(unsigned kinds) { seed_kinds_ = kinds; }
  void SetConcurrentExtensions(bool two) override;
  const std::vector<uint8_t> &LastConcatenatedQueries() const { return concat_; }

  // PresearchDevice
  int BeginQueryChunk(const QuerySet &q, const std::vector<uint8_t> &concat,
                      const std::vector<uint64_t> &coff, const std::vector<int> &cutoffs,
                      const std::vector<int> &triggers, int presearch_cutoff,
                      AlignStats *st) override;
  int SeedHits(size_t chunk, uint32_t qb, uint32_t qe, SeedHitView *hits,
               AlignStats *st) override;
  int GappedExtend(size_t chunk, uint32_t n_reverse, uint32_t n_forward, uint32_t *qpos,
                   uint32_t *dpos, int *scores, uint32_t *cells, const uint16_t *keys,
                   bool learned, AlignStats *st, const uint32_t *bucket_off) override;
  int GappedPrepare(int slot, size_t chunk, uint32_t n_reverse, uint32_t n_forward,
                    const uint32_t *qpos, const uint32_t *dpos, const uint16_t *keys,
                    bool learned, const uint32_t *bucket_off) override;
  int GappedRun(int slot, size_t chunk, uint32_t n_reverse, uint32_t n_forward, uint32_t *qpos,
                uint32_t *dpos, int *scores, uint32_t *cells, AlignStats *st,
                const uint32_t *bucket_off) override;
  void GraphCounts(uint64_t *instantiations, uint64_t *updates) const {
    *instantiations = graph_instantiations_;
    *updates = graph_updates_;
  }
  GappedTuner Tuner() {
    std::lock_guard<std::mutex> lock(tuner_mutex_);
    return tuner_;
  }
  int Tracebacks(size_t chunk, const std::vector<TracebackJob> &jobs, int cutoff,
                 std::vector<TracebackOut> *out, AlignStats *st) override;
  void EndQueryChunk() override;
  uint8_t *JobBuffer(size_t slot, size_t bytes) override;

 private:
  struct DeviceChunk;
  AlignerOptions opt_;
  const Database *db_ = nullptr;
  std::vector<std::unique_ptr<DeviceChunk> > chunks_;
  hipStream_t stream_ = nullptr;         // uploads and the seed stage
  // result build (tracebacks): its own stream, since the batched form runs
  // it on another thread while the extension stream may be capturing a graph
  hipStream_t traceback_stream_ = nullptr;
  int *d_matrix_ = nullptr;
  uint8_t *d_reduced_ = nullptr;
  SeedStage seed_;
  DeviceListBuilder lists_;  // query hash lists of the resident chunk
  // per pipeline slot: the gapped extender (its workspace) and the
  // extension stream, so that two tasks' extensions can run at once
  gz_gapped_extender_s *gapped_[2] = {nullptr, nullptr};
  std::vector<GappedTrace> traces_;
  std::vector<uint8_t> concat_;
  // resident query chunk
  DeviceQueryLists dq_;
  SeedStageParams sp_;
  static constexpr size_t kNoChunk = ~(size_t)0;
  size_t lookup_chunk_ = kNoChunk;  // database chunk of the seed stage's lookup
  std::vector<void *> query_buffers_;
  unsigned seed_kinds_ = 3;
  // timing events: [0, 2) the gapped kernels', [4, 6) the result build's
  // (the stages may run on different host threads)
  hipEvent_t ev_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // The gapped stage's two pipeline slots: GappedPrepare uploads a slot's
  // jobs and orders them on the schedule stream, GappedRun extends them on
  // the extension stream once the slot's `ready` event has passed, so one
  // slot's ordering overlaps the other's extension.
  struct GappedSlot {
    void *ws = nullptr;  // positions, scores, cells, keys, permutation, scratch
    size_t cap = 0;
    size_t n = 0;
    bool order = false, gated = false;
    bool whole = false;  // one launch per side (GappedTuner::PickLayout)
    uint32_t *pq = nullptr, *pd = nullptr, *cells = nullptr, *perm = nullptr;
    int *sc = nullptr;
    uint32_t *counts = nullptr;  // pinned: the two sides' gated job counts
    uint32_t *d_counts = nullptr;  // the same on the device
    hipEvent_t ready = nullptr, ev_begin = nullptr, ev_end = nullptr;
  };
  GappedSlot slots_[2];
  hipStream_t schedule_stream_ = nullptr;
  // the extension streams: [0] always, [1] created for concurrent
  // extensions (each slot then extends on its own stream and extender)
  hipStream_t extend_streams_[2] = {nullptr, nullptr};
  bool two_extensions_ = false;
  hipEvent_t run_ev_[2][2] = {};  // per slot: around its gapped kernels
  std::mutex stats_mutex_;        // AlignStats updates from both slots
  // result-build workspace
  void *tb_ws_ = nullptr;
  size_t tb_ws_cap_ = 0;
  uint32_t max_query_length_ = 0;
  // pinned job buffers per pipeline slot
  std::vector<std::pair<void *, size_t> > pinned_jobs_;
  // the learned schedule's gate (query side residues)
  uint32_t learned_gate_ = 44;
  GappedTuner tuner_;
  std::mutex tuner_mutex_;
  hipEvent_t tune_ev_[2][2 * kLengthBuckets] = {};
  // per (bucket, configuration): the executable graph of its launch
  // sequence, updated with each call's arguments
  bool use_graphs_ = true;
  int forced_layout_ = -1;  // GappedTuner layout from GZ_GAPPED_LAYOUT (-1: tuned)
  hipGraphExec_t graph_exec_[2 * kLengthBuckets * GappedTuner::kConfigs] = {};
  uint64_t graph_updates_ = 0, graph_instantiations_ = 0;
  hipError_t LaunchCaptured(int slot, hipStream_t s, const std::function<bool()> &launch);
};

}  // namespace gz
