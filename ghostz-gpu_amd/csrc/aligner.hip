// aligner.hip -- see aligner.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <mutex>

#include "aligner.h"
#include "cost_model_gpu.h"
#include "gz_scan.h"
#include "schedule.h"
#include "traceback.h"
#include "ungapped.h"
#include "cost_model.h"
#include "gz_runtime.h"
#include "gz_common.h"
#include "query_lists.h"

namespace gz {

namespace {

template <class T>
hipError_t Upload(T **d, const std::vector<T> &h, hipStream_t s) {
  hipError_t e = gz::Malloc((void **)d, std::max<size_t>(h.size() * sizeof(T), 16));
  if (e != hipSuccess) return e;
  if (!h.empty()) e = hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
  return e;
}

// packs host codes through a temporary device copy
hipError_t UploadPacked(uint64_t **d, const std::vector<uint8_t> &codes, hipStream_t s) {
  uint8_t *tmp = nullptr;
  hipError_t e = gz::Malloc((void **)&tmp, std::max<size_t>(codes.size(), 16));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(tmp, codes.data(), codes.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = gz::Malloc((void **)d, PackedWords(codes.size()) * 8);
  if (e == hipSuccess) {
    gz_stream_s wrap{s};
    if (gz_pack_sequence_async(tmp, codes.size(), *d, &wrap) != GZ_OK) e = hipErrorUnknown;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)gz::Free(tmp);
  return e;
}

}  // namespace

#define GZ_TRY(x)                      \
  do {                                 \
    hipError_t _e = (x);               \
    if (_e != hipSuccess) return _e;   \
  } while (0)

struct Aligner::DeviceChunk {
  uint64_t *packed = nullptr;
  uint32_t *hashes = nullptr, *rep_off = nullptr, *rep_pos = nullptr, *mem_off = nullptr,
           *mem_pos = nullptr, *val_off = nullptr, *val_pos = nullptr;
  uint32_t nhash = 0;
  ~DeviceChunk() {
    void *ps[] = {packed, hashes, rep_off, rep_pos, mem_off, mem_pos, val_off, val_pos};
    for (void *p : ps)
      if (p) (void)gz::Free(p);
  }
};

Aligner::Aligner(const AlignerOptions &opt) : opt_(opt) {
  // jobs whose query side has at least this many residues get a prediction
  // (GZ_LEARNED_GATE overrides the default, for tuning)
  const char *g = getenv("GZ_LEARNED_GATE");
  if (g && *g) learned_gate_ = (uint32_t)strtoul(g, nullptr, 10);
  // the batched form launches each bucket's gapped sequence as a graph
  // (GZ_NO_GRAPHS=1 launches the same kernels directly, for comparison)
  const char *ng = getenv("GZ_NO_GRAPHS");
  use_graphs_ = !(ng && *ng && *ng != '0');
  // the batched form's gapped layout per task is tuned (GZ_GAPPED_LAYOUT=
  // buckets or whole fixes it, for comparison and tests)
  const char *gl = getenv("GZ_GAPPED_LAYOUT");
  if (gl && strcmp(gl, "buckets") == 0) forced_layout_ = 0;
  if (gl && strcmp(gl, "whole") == 0) forced_layout_ = 1;
}

Aligner::~Aligner() {
  EndQueryChunk();
  chunks_.clear();
  for (GappedSlot &S : slots_) {
    if (S.ws) (void)gz::Free(S.ws);
    if (S.counts) (void)gz::HostFree(S.counts);
    for (hipEvent_t e : {S.ready, S.ev_begin, S.ev_end})
      if (e) (void)hipEventDestroy(e);
  }
  if (tb_ws_) (void)gz::Free(tb_ws_);
  for (auto &b : pinned_jobs_)
    if (b.first) (void)gz::HostFree(b.first);
  for (hipEvent_t e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto &slot : tune_ev_)
    for (hipEvent_t e : slot)
      if (e) (void)hipEventDestroy(e);
  for (auto &slot : run_ev_)
    for (hipEvent_t e : slot)
      if (e) (void)hipEventDestroy(e);
  for (hipGraphExec_t x : graph_exec_)
    if (x) (void)hipGraphExecDestroy(x);
  for (gz_gapped_extender_s *g : gapped_)
    if (g) gz_gapped_extender_destroy(g);
  if (d_matrix_) (void)gz::Free(d_matrix_);
  if (d_reduced_) (void)gz::Free(d_reduced_);
  if (stream_) (void)hipStreamDestroy(stream_);
  for (hipStream_t x : extend_streams_)
    if (x) (void)hipStreamDestroy(x);
  if (traceback_stream_) (void)hipStreamDestroy(traceback_stream_);
  if (schedule_stream_) (void)hipStreamDestroy(schedule_stream_);
}

hipError_t Aligner::SetDatabase(const Database *db) {
  db_ = db;
  if (!stream_) GZ_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  if (!extend_streams_[0])
    GZ_TRY(hipStreamCreateWithFlags(&extend_streams_[0], hipStreamNonBlocking));
  if (!traceback_stream_)
    GZ_TRY(hipStreamCreateWithFlags(&traceback_stream_, hipStreamNonBlocking));
  if (!schedule_stream_)
    GZ_TRY(hipStreamCreateWithFlags(&schedule_stream_, hipStreamNonBlocking));
  chunks_.clear();
  for (const DbChunk &c : db->chunks) {
    std::unique_ptr<DeviceChunk> d(new DeviceChunk());
    GZ_TRY(UploadPacked(&d->packed, c.seq, stream_));
    GZ_TRY(Upload(&d->hashes, c.hashes, stream_));
    GZ_TRY(Upload(&d->rep_off, c.rep_off, stream_));
    GZ_TRY(Upload(&d->rep_pos, c.rep_pos, stream_));
    GZ_TRY(Upload(&d->mem_off, c.mem_off, stream_));
    GZ_TRY(Upload(&d->mem_pos, c.mem_pos, stream_));
    GZ_TRY(Upload(&d->val_off, c.val_off, stream_));
    GZ_TRY(Upload(&d->val_pos, c.val_pos, stream_));
    d->nhash = (uint32_t)c.hashes.size();
    chunks_.push_back(std::move(d));
  }
  std::vector<int> m(opt_.search.matrix, opt_.search.matrix + opt_.search.nl * opt_.search.nl);
  if (!d_matrix_) GZ_TRY(Upload(&d_matrix_, m, stream_));
  std::vector<uint8_t> rm(db->reduced_map, db->reduced_map + 25);
  if (d_reduced_) GZ_TRY(gz::Free(d_reduced_));
  GZ_TRY(Upload(&d_reduced_, rm, stream_));
  GZ_TRY(hipStreamSynchronize(stream_));
  // the batched form runs the first ring as a persistent work-queue launch
  for (gz_gapped_extender_s *&g : gapped_) {
    if (!g && gz_gapped_extender_create(&g) != GZ_OK) return hipErrorUnknown;
    if (gz_gapped_extender_set_persistent(g, opt_.batched ? 1 : 0) != GZ_OK)
      return hipErrorUnknown;
  }
  return hipSuccess;
}

hipError_t Aligner::Align(const QuerySet &queries, std::string *output, AlignStats *st) {
  if (!db_) return hipErrorInvalidValue;
  const int r = AlignQueries(queries, *db_, opt_, this, output, st, &traces_, &concat_);
  EndQueryChunk();
  return r == 0 ? hipSuccess : (hipError_t)r;
}

int Aligner::BeginQueryChunk(const QuerySet &q, const std::vector<uint8_t> &concat,
                             const std::vector<uint64_t> &coff, const std::vector<int> &cut,
                             const std::vector<int> &trig, int presearch_cutoff,
                             AlignStats *st) {
  EndQueryChunk();
  // the seed walk's stop scores serve cutoffs in [1, kMaxUngappedCutoff]
  // (byte tables: [1, kMaxByteCutoff])
  bool byte_tables = true;
  for (int c : cut) {
    if (c < 1 || c > kMaxUngappedCutoff) return (int)hipErrorInvalidValue;
    byte_tables = byte_tables && c <= kMaxByteCutoff;
  }
  for (uint32_t i = 0; i < opt_.search.nl * opt_.search.nl; ++i) {
    const int v = opt_.search.matrix[i];
    if (v < -kMaxUngappedCutoff + 1 || v > kMaxUngappedCutoff - 1) return (int)hipErrorInvalidValue;
    byte_tables = byte_tables && v >= -127 && v <= 127;
  }
  max_query_length_ = q.max_length;
  hipStream_t s = stream_;
  uint64_t *d_q = nullptr;
  int *d_cut = nullptr, *d_trig = nullptr;
  hipError_t e = UploadPacked(&d_q, concat, s);
  query_buffers_.push_back(d_q);
  if (e == hipSuccess) e = Upload(&d_cut, cut, s);
  query_buffers_.push_back(d_cut);
  if (e == hipSuccess) e = Upload(&d_trig, trig, s);
  query_buffers_.push_back(d_trig);
  if (e != hipSuccess) return (int)e;
  // the hash lists, built on the device from the packed concatenation
  dq_ = DeviceQueryLists();
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t entries = 0;
  e = lists_.Build(d_q, concat.size(), q, coff, db_->hash, db_->subsequence_length, s, &dq_,
                   &entries);
  if (e != hipSuccess) return (int)e;
  st->seconds_query_lists +=
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  st->lists += dq_.nlists;
  dq_.packed = d_q;
  dq_.cutoffs = d_cut;
  dq_.triggers = d_trig;
  dq_.nqueries = (uint32_t)q.size();
  e = seed_.PrepareEntries(&dq_, entries, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return (int)e;
  lookup_chunk_ = kNoChunk;
  sp_ = SeedStageParams();
  sp_.matrix = d_matrix_;
  sp_.nl = opt_.search.nl;
  sp_.delimiter = kDelimiter;
  sp_.reduced_map = d_reduced_;
  sp_.subsequence_length = db_->subsequence_length;
  {
    const float rate = 0.8f;
    const uint32_t L = db_->subsequence_length;
    const char threshold = L - rate * L;  // SeedSearcherGpu::Reset (:58-69)
    sp_.distance_threshold = threshold;
  }
  sp_.max_mismatches = db_->max_mismatches;
  sp_.byte_tables = byte_tables;

  for (gz_gapped_extender_s *g : gapped_)
    if (gz_gapped_extender_set_max_sequence_length(g, q.max_length) != GZ_OK ||
        gz_gapped_extender_set_queries(g, kDelimiter, d_q) != GZ_OK ||
        gz_gapped_extender_set_score_parameters(g, d_matrix_, opt_.search.nl,
                                                opt_.search.gap_open, opt_.search.gap_extension,
                                                presearch_cutoff) != GZ_OK)
      return (int)hipErrorUnknown;
  return 0;
}

void Aligner::EndQueryChunk() {
  for (void *p : query_buffers_)
    if (p) (void)gz::Free(p);
  query_buffers_.clear();
  dq_ = DeviceQueryLists();
}

int Aligner::SeedHits(size_t ci, uint32_t qb, uint32_t qe, SeedHitView *hits,
                      AlignStats *st) {
  const DeviceChunk &dc = *chunks_[ci];
  DeviceDbChunk ddb;
  ddb.packed = dc.packed;
  ddb.nhash = dc.nhash;
  ddb.hashes = dc.hashes;
  ddb.rep_off = dc.rep_off;
  ddb.rep_pos = dc.rep_pos;
  ddb.mem_off = dc.mem_off;
  ddb.mem_pos = dc.mem_pos;
  ddb.val_off = dc.val_off;
  ddb.val_pos = dc.val_pos;
  sp_.clustering = db_->chunks[ci].clustering;
  if (lookup_chunk_ != ci) {
    lookup_chunk_ = kNoChunk;
    const hipError_t e = seed_.Lookup(dq_, ddb, sp_.clustering, stream_);
    if (e != hipSuccess) return (int)e;
    lookup_chunk_ = ci;
  }
  SeedStageStats ss;
  const hipError_t e =
      seed_.Run(dq_, ddb, sp_, qb, qe, stream_, hits, &ss, seed_kinds_, /*compact=*/true);
  if (e != hipSuccess) return (int)e;
  st->units_rep += ss.units_rep;
  st->units_value += ss.units_value;
  for (int k = 0; k < 2; ++k) {
    st->seed_kernel_ms[k] += ss.kernel_ms[k];
    st->seed_launches[k] += ss.launches[k];
  }
  return 0;
}

int GappedTuner::Pick(int bucket) const {
  for (int c = 0; c < kConfigs; ++c)
    if (trials[bucket][c] < kTrials) return c;
  int best = 0;
  for (int c = 1; c < kConfigs; ++c)
    if (ms[bucket][c] * jobs[bucket][best] < ms[bucket][best] * jobs[bucket][c]) best = c;
  return best;
}

void GappedTuner::Record(int bucket, int config, double kernel_ms, double njobs) {
  ms[bucket][config] += kernel_ms;
  jobs[bucket][config] += njobs;
  ++trials[bucket][config];
}

void Aligner::SetConcurrentExtensions(bool two) {
  two_extensions_ = two;
  // the second extension stream exists from here on (created before the
  // pipeline's threads start)
  if (two && !extend_streams_[1] &&
      hipStreamCreateWithFlags(&extend_streams_[1], hipStreamNonBlocking) != hipSuccess)
    extend_streams_[1] = nullptr;  // GappedRun creates it then, or reports the error
}

int GappedTuner::PickLayout() const {
  if (layout_trials[0] < kTrials) return 0;
  if (layout_trials[1] < kTrials) return 1;
  return layout_ms[1] * layout_jobs[0] < layout_ms[0] * layout_jobs[1] ? 1 : 0;
}

void GappedTuner::RecordLayout(int layout, double kernel_ms, double njobs) {
  layout_ms[layout] += kernel_ms;
  layout_jobs[layout] += njobs;
  ++layout_trials[layout];
}

int Aligner::GappedExtend(size_t ci, uint32_t nr, uint32_t nf, uint32_t *qpos, uint32_t *dpos,
                          int *scores, uint32_t *cells, const uint16_t *keys, bool learned,
                          AlignStats *st, const uint32_t *bucket_off) {
  const int r = GappedPrepare(0, ci, nr, nf, qpos, dpos, keys, learned, bucket_off);
  if (r != 0) return r;
  return GappedRun(0, ci, nr, nf, qpos, dpos, scores, cells, st, bucket_off);
}

int Aligner::GappedPrepare(int slot, size_t ci, uint32_t nr, uint32_t nf, const uint32_t *qpos,
                           const uint32_t *dpos, const uint16_t *keys, bool learned,
                           const uint32_t *bucket_off) {
  GappedSlot &S = slots_[slot];
  S.n = (size_t)nr + nf;
  S.order = false;
  S.gated = false;
  S.whole = false;
  if (S.n == 0) return 0;
  if (bucket_off && nr == nf) {
    std::lock_guard<std::mutex> lock(tuner_mutex_);
    S.whole = forced_layout_ >= 0 ? forced_layout_ == 1 : tuner_.PickLayout() == 1;
  }
  const size_t n = S.n;
  hipStream_t s = schedule_stream_;
  if (!S.ready) {
    GZ_TRY(hipEventCreateWithFlags(&S.ready, hipEventDisableTiming));
    GZ_TRY(hipEventCreate(&S.ev_begin));
    GZ_TRY(hipEventCreate(&S.ev_end));
  }
  if (!S.counts) GZ_TRY(gz::HostMalloc(&S.counts, 64));
  const DeviceChunk &dc = *chunks_[ci];
  const GpuCostModel *model =
      learned ? dynamic_cast<const GpuCostModel *>(opt_.cost_model) : nullptr;
  S.order = keys != nullptr || model != nullptr;
  // workspace: positions, scores, cells; and for an ordered run the keys,
  // the permutation, predicted costs and the sort scratch (the extension
  // kernels read the jobs through the permutation)
  const size_t a4 = AlignUp(n * 4, 256), a2 = AlignUp(n * 2, 256);
  // the learned schedule with host lengths: the model orders the gated jobs
  // inside their length bins (schedule.h)
  const bool gated = model != nullptr && keys != nullptr;
  size_t bytes = 4 * a4;
  if (S.order) bytes += a2 + 2 * a4 + AlignUp(ScheduleTempBytes(n), 256);
  if (gated) bytes += a2 + a4 + 256;
  if (S.cap < bytes) {
    if (S.ws) (void)gz::Free(S.ws);
    S.ws = nullptr;
    S.cap = 0;
    GZ_TRY(gz::Malloc(&S.ws, bytes));
    S.cap = bytes;
  }
  char *w = (char *)S.ws;
  auto take = [&](size_t b) {
    char *p = w;
    w += b;
    return p;
  };
  S.pq = (uint32_t *)take(a4);
  S.pd = (uint32_t *)take(a4);
  S.sc = (int *)take(a4);
  S.cells = (uint32_t *)take(a4);
  uint16_t *d_keys = nullptr;
  float *d_pred = nullptr;
  void *d_tmp = nullptr;
  S.perm = nullptr;
  if (S.order) {
    d_keys = (uint16_t *)take(a2);
    S.perm = (uint32_t *)take(a4);
    d_pred = (float *)take(a4);
    d_tmp = take(AlignUp(ScheduleTempBytes(n), 256));
  }
  uint16_t *d_len = nullptr;
  uint32_t *d_index = nullptr, *d_count = nullptr;
  if (gated) {
    d_len = (uint16_t *)take(a2);
    d_index = (uint32_t *)take(a4);
    d_count = (uint32_t *)take(256);
  }
  hipError_t e = hipMemcpyAsync(S.pq, qpos, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(S.pd, dpos, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && S.order) {
    // order the jobs by descending key inside each side, on the device
    e = hipEventRecord(S.ev_begin, s);
    if (e == hipSuccess && gated) {
      // length bins for every job, the model's prediction bins for the
      // jobs whose side may run long
      e = hipMemcpyAsync(d_len, keys, n * 2, hipMemcpyHostToDevice, s);
      // the gated counts stay on the device: the launches are sized for
      // every job of a side and read the count (no wait in the stream)
      const uint32_t off[2] = {0, nr}, cnt[2] = {nr, nf};
      for (int k = 0; k < 2 && e == hipSuccess; ++k)
        e = BinAndGateJobs(d_len + off[k], cnt[k], learned_gate_, d_keys + off[k],
                           d_index + off[k], d_count + k, s);
      for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = model->PredictIndexedOnDevice(dq_.packed, dc.packed, S.pq + off[k], S.pd + off[k],
                                          d_index + off[k], cnt[k], k == 0, d_pred + off[k], s,
                                          d_count + k);
        if (e == hipSuccess)
          e = AddPredictionBins(d_pred + off[k], d_index + off[k], cnt[k], d_keys + off[k], s,
                                d_count + k);
      }
      // the counts come back with the extension's results
      S.d_counts = d_count;
      S.gated = true;
    } else if (e == hipSuccess && model) {
      // the host lays out job nr + k as the forward side of job k (both
      // positions one later), so one feature pass serves both sides; were
      // that not so, only the order (never a result) would differ
      if (nr == nf)
        e = model->PredictBothSidesOnDevice(dq_.packed, dc.packed, S.pq, S.pd, nr, d_pred,
                                            d_pred + nr, s);
      else
        e = model->PredictOnDevice(dq_.packed, dc.packed, S.pq, S.pd, nr, true, d_pred, s);
      if (e == hipSuccess && nr != nf)
        e = model->PredictOnDevice(dq_.packed, dc.packed, S.pq + nr, S.pd + nr, nf, false,
                                   d_pred + nr, s);
      // log2 cells in steps of 1/32
      if (e == hipSuccess) e = CostToKey(d_pred, (uint32_t)n, 32.f, d_keys, s);
    } else if (e == hipSuccess) {
      e = hipMemcpyAsync(d_keys, keys, n * 2, hipMemcpyHostToDevice, s);
    }
    if (bucket_off && !S.whole) {
      // each side's buckets are ordered separately (positions in the
      // permutation are bucket-relative)
      for (int side = 0; side < 2 && e == hipSuccess; ++side)
        for (int b = 0; b < kLengthBuckets && e == hipSuccess; ++b) {
          const size_t o = (side ? nr : 0) + bucket_off[b];
          e = DeviceOrderByKey(d_keys + o, bucket_off[b + 1] - bucket_off[b], S.perm + o, d_tmp,
                               s);
        }
    } else {
      if (e == hipSuccess) e = DeviceOrderByKey(d_keys, nr, S.perm, d_tmp, s);
      if (e == hipSuccess) e = DeviceOrderByKey(d_keys + nr, nf, S.perm + nr, d_tmp, s);
    }
    if (e == hipSuccess) e = hipEventRecord(S.ev_end, s);
  }
  // the extension stream waits for this event before it reads the slot
  if (e == hipSuccess) e = hipEventRecord(S.ready, s);
  return e == hipSuccess ? 0 : (int)e;
}

int Aligner::GappedRun(int slot, size_t ci, uint32_t nr, uint32_t nf, uint32_t *qpos,
                       uint32_t *dpos, int *scores, uint32_t *cells, AlignStats *st,
                       const uint32_t *bucket_off) {
  GappedSlot &S = slots_[slot];
  const size_t n = S.n;
  if (n == 0) return 0;
  if (n != (size_t)nr + nf) return (int)hipErrorInvalidValue;
  // with concurrent extensions each slot extends on its own stream with its
  // own extender, so the other slot's task can run at the same time
  const int x = two_extensions_ ? slot : 0;
  if (!extend_streams_[x])
    GZ_TRY(hipStreamCreateWithFlags(&extend_streams_[x], hipStreamNonBlocking));
  hipStream_t s = extend_streams_[x];
  gz_gapped_extender_s *gapped = gapped_[x];
  hipEvent_t *ev = run_ev_[x];
  hipEvent_t *tune_ev = tune_ev_[x];
  for (int i = 0; i < 2; ++i)
    if (!ev[i]) GZ_TRY(hipEventCreate(&ev[i]));
  const DeviceChunk &dc = *chunks_[ci];
  if (gz_gapped_extender_set_database(gapped, dc.packed) != GZ_OK) return (int)hipErrorUnknown;
  const bool order = S.order;
  uint32_t *rc = cells ? S.cells : nullptr;
  hipError_t e = hipStreamWaitEvent(s, S.ready, 0);
  // each side's jobs run in the permutation's order (when ordered) and
  // leave their results at their own indices
  gz_stream_s wrap{s};
  auto side = [&](size_t off, size_t count, int reverse) {
    uint32_t *c = rc ? rc + off : nullptr;
    const gz_status r =
        order ? gz_gapped_extender_extend_one_side_score_only_ordered_async(
                    gapped, count, reverse, S.perm + off, S.pq + off, S.pd + off, S.sc + off,
                    c, &wrap)
              : gz_gapped_extender_extend_one_side_score_only_cells_async(
                    gapped, count, reverse, S.pq + off, S.pd + off, S.sc + off, c, &wrap);
    return r == GZ_OK;
  };
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  int tuned[kLengthBuckets] = {-1, -1, -1, -1};
  bool settled = true;  // every bucket ran its chosen configuration
  const bool bucketed = bucket_off && nr == nf && !S.whole;
  if (bucketed && e == hipSuccess) {
    // the workspace for the largest bucket, allocated before any capture
    uint32_t most = 0;
    for (int b = 0; b < kLengthBuckets; ++b) most = std::max(most, bucket_off[b + 1] - bucket_off[b]);
    if (gz_gapped_extender_reserve(gapped, most, &wrap) != GZ_OK) e = hipErrorUnknown;
  }
  if (bucketed) {
    // every query-length bucket with the configuration the tuner picks for
    // it (first ring width, persistent work queue or static grid), each
    // bucket timed on its own
    for (int b = 0; b < kLengthBuckets && e == hipSuccess; ++b) {
      const uint32_t o = bucket_off[b], c = bucket_off[b + 1] - o;
      if (c == 0) continue;
      if (!tune_ev[2 * b]) {
        e = hipEventCreate(&tune_ev[2 * b]);
        if (e == hipSuccess) e = hipEventCreate(&tune_ev[2 * b + 1]);
        if (e != hipSuccess) break;
      }
      std::lock_guard<std::mutex> lock(tuner_mutex_);
      tuned[b] = tuner_.Pick(b);
      for (int c = 0; c < GappedTuner::kConfigs; ++c)
        if (tuner_.trials[b][c] < GappedTuner::kTrials) settled = false;
      if (gz_gapped_extender_set_narrow_ring(gapped, GappedTuner::Narrow(tuned[b])) != GZ_OK ||
          gz_gapped_extender_set_persistent(gapped, GappedTuner::Persistent(tuned[b])) != GZ_OK)
        e = hipErrorUnknown;
      if (e == hipSuccess) e = hipEventRecord(tune_ev[2 * b], s);
      // the bucket's launch sequence (both sides: counter resets, the
      // first ring, the 64-column ring, the global ring) as one graph
      auto both = [&]() { return side(o, c, 1) && side(nr + o, c, 0); };
      if (e == hipSuccess) {
        const int key = (x * kLengthBuckets + b) * GappedTuner::kConfigs + tuned[b];
        if (use_graphs_) e = LaunchCaptured(key, s, both);
        else if (!both()) e = hipErrorUnknown;
      }
      if (e == hipSuccess) e = hipEventRecord(tune_ev[2 * b + 1], s);
    }
  } else {
    if (S.whole &&
        (gz_gapped_extender_set_narrow_ring(
             gapped, GappedTuner::Narrow(GappedTuner::kWholeConfig)) != GZ_OK ||
         gz_gapped_extender_set_persistent(
             gapped, GappedTuner::Persistent(GappedTuner::kWholeConfig)) != GZ_OK))
      e = hipErrorUnknown;
    if (e == hipSuccess && nr > 0 && !side(0, nr, 1)) e = hipErrorUnknown;
    if (e == hipSuccess && nf > 0 && !side(nr, nf, 0)) e = hipErrorUnknown;
  }
  if (e == hipSuccess) e = hipEventRecord(ev[1], s);
  // the copies back are queued once the kernels are done: a device-to-host
  // copy queued behind a running kernel holds up the other streams' copies
  // (the seed stage's hits, the chain filter's records) until that kernel
  // ends
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  if (e == hipSuccess && S.gated)
    e = hipMemcpyAsync(S.counts, S.d_counts, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(qpos, S.pq, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dpos, S.pd, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(scores, S.sc, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && cells) e = hipMemcpyAsync(cells, S.cells, n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  // the slot's schedule ran before its ready event, which this stream waited for
  if (e == hipSuccess && S.gated) {
    std::lock_guard<std::mutex> lock(stats_mutex_);
    st->gated_jobs += (uint64_t)S.counts[0] + S.counts[1];
  }
  uint32_t fails = 0;
  if (e == hipSuccess && (gz_gapped_extender_failures(gapped, &fails) != GZ_OK || fails != 0))
    e = hipErrorUnknown;
  float ms = 0, ms2 = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
  if (e == hipSuccess && order) e = hipEventElapsedTime(&ms2, S.ev_begin, S.ev_end);
  for (int b = 0; b < kLengthBuckets && e == hipSuccess; ++b) {
    if (tuned[b] < 0) continue;
    float bm = 0;
    e = hipEventElapsedTime(&bm, tune_ev[2 * b], tune_ev[2 * b + 1]);
    if (e == hipSuccess) {
      std::lock_guard<std::mutex> lock(tuner_mutex_);
      tuner_.Record(b, tuned[b], bm, 2.0 * (bucket_off[b + 1] - bucket_off[b]));
    }
  }
  if (e == hipSuccess && bucket_off && nr == nf && (S.whole || settled)) {
    std::lock_guard<std::mutex> lock(tuner_mutex_);
    tuner_.RecordLayout(S.whole ? 1 : 0, ms, (double)n);
  }
  if (e != hipSuccess) return (int)e;
  std::lock_guard<std::mutex> lock(stats_mutex_);
  st->gapped_kernel_ms += ms;
  st->schedule_kernel_ms += ms2;
  st->gapped_launches += (nr > 0) + (nf > 0);
  return 0;
}

int Aligner::Tracebacks(size_t ci, const std::vector<TracebackJob> &jobs, int cutoff,
                        std::vector<TracebackOut> *out, AlignStats *st) {
  const size_t n = jobs.size();
  out->resize(n);
  if (n == 0) return 0;
  hipStream_t s = traceback_stream_;
  for (int i = 4; i < 6; ++i)
    if (!ev_[i]) GZ_TRY(hipEventCreate(&ev_[i]));
  // m jobs with the given per-lane capacities: lanes resident at once (up
  // to seven 64-lane blocks per CU), bounded by a 12 GB scratch budget
  auto pass = [&](const TracebackJob *hjobs, size_t m, const LaneCaps &caps,
                  TracebackOut *hout) -> hipError_t {
    const size_t budget = (size_t)12 << 30;
    size_t lanes = std::min<size_t>((m + 63) / 64 * 64, (size_t)7 * 64 * 256);
    lanes = std::max<size_t>(64, std::min(lanes, budget / caps.bytes_per_lane / 64 * 64));
    const size_t a_jobs = AlignUp(m * sizeof(TracebackJob), 256);
    const size_t a_outs = AlignUp(m * sizeof(TracebackOut), 256);
    const size_t bytes = a_jobs + a_outs + lanes * caps.bytes_per_lane;
    if (tb_ws_cap_ < bytes) {
      if (tb_ws_) (void)gz::Free(tb_ws_);
      tb_ws_ = nullptr;
      tb_ws_cap_ = 0;
      const hipError_t a = gz::Malloc(&tb_ws_, bytes);
      if (a != hipSuccess) return a;
      tb_ws_cap_ = bytes;
    }
    TracebackJob *d_jobs = (TracebackJob *)tb_ws_;
    TracebackOut *d_outs = (TracebackOut *)((char *)tb_ws_ + a_jobs);
    uint8_t *scratch = (uint8_t *)tb_ws_ + a_jobs + a_outs;
    hipError_t e = hipMemcpyAsync(d_jobs, hjobs, m * sizeof(TracebackJob),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ev_[4], s);
    if (e == hipSuccess)
      e = TracebackLaunch(dq_.packed, chunks_[ci]->packed, d_jobs, (uint32_t)m, d_matrix_,
                          opt_.search.nl, opt_.search.gap_open, opt_.search.gap_extension,
                          cutoff, scratch, caps, (uint32_t)lanes, d_outs, s);
    if (e == hipSuccess) e = hipEventRecord(ev_[5], s);
    // the outputs' copy is queued once the kernel is done (as in GappedRun)
    if (e == hipSuccess) e = hipEventSynchronize(ev_[5]);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hout, d_outs, m * sizeof(TracebackOut), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev_[4], ev_[5]);
    if (e == hipSuccess) st->traceback_kernel_ms += ms;
    return e;
  };
  const LaneCaps caps = TracebackCaps(max_query_length_, db_->max_length);
  GZ_TRY(pass(jobs.data(), n, caps, out->data()));
  // results whose op rows outgrow the lane's arena are rebuilt on the host
  // (on all threads): a long band on one lane runs far slower than on a core
  return 0;
}

hipError_t Aligner::LaunchCaptured(int slot, hipStream_t s, const std::function<bool()> &launch) {
  hipGraph_t g = nullptr;
  bool ok = false;
  hipError_t ec = hipSuccess;
  {
    // no other thread of the library frees or allocates while the capture runs
    std::lock_guard<std::recursive_mutex> lock(CaptureLock());
    GZ_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    ok = launch();
    ec = hipStreamEndCapture(s, &g);
  }
  if (!ok || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return ok ? ec : hipErrorUnknown;
  }
  // the slot's executable graph takes the new arguments when the launch
  // sequence has the same shape, else it is instantiated again
  hipGraphExec_t &x = graph_exec_[slot];
  hipError_t e = hipSuccess;
  if (x) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult r = hipGraphExecUpdateError;
    if (hipGraphExecUpdate(x, g, &bad, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    } else {
      ++graph_updates_;
    }
  }
  if (!x) {
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e == hipSuccess) ++graph_instantiations_;
  }
  if (e == hipSuccess) e = hipGraphLaunch(x, s);
  (void)hipGraphDestroy(g);
  return e;
}


uint8_t *Aligner::JobBuffer(size_t slot, size_t bytes) {
  if (pinned_jobs_.size() <= slot) pinned_jobs_.resize(slot + 1, std::make_pair(nullptr, 0));
  std::pair<void *, size_t> &b = pinned_jobs_[slot];
  if (b.second < bytes) {
    if (b.first) (void)gz::HostFree(b.first);
    b.first = nullptr;
    b.second = 0;
    const size_t want = bytes + bytes / 4;
    if (gz::HostMalloc(&b.first, want, hipHostMallocDefault) != hipSuccess) {
      b.first = nullptr;
      return PresearchDevice::JobBuffer(slot, bytes);  // pageable fallback
    }
    b.second = want;
  }
  return (uint8_t *)b.first;
}

}  // namespace gz
