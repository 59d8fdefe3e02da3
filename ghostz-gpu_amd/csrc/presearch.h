// presearch.h -- the host side of AlignerGpu's search (src/aligner_gpu.cpp,
// src/aligner_gpu_presearch_thread.cpp:140-329): query chunking, hash
// lists, per-query seed order, chain filter, gapped-extension schedule,
// AddResults, then BuildResults and WriteOutput.  The device stages (seed
// search and one-sided gapped extension) are reached through
// PresearchDevice; the product implements it with the HIP kernels
// (aligner.hip).
#pragma once

#include <stdint.h>

#include <functional>

#include <string>
#include <vector>

#include "database.h"
#include "queries.h"
#include "query_lists.h"
#include "results.h"

namespace gz {

// One passing seed of the device seed stage with the key (list rank,
// representative or value index, member index + 1 or 0, entry index) that
// orders it exactly as the one-thread CPU search enumerates it.
struct SeedHit {
  uint32_t qid;
  uint32_t qpos;   // concatenated query position
  uint32_t dpos;   // database chunk position
  uint32_t list;   // list rank
  uint32_t rv;     // representative index, or number of reps + value index
  uint32_t member; // 0 for a representative / value, 1 + member index
  uint32_t entry;  // entry index within the list
};

// A passing seed's query and database positions.
struct HitPosition {
  uint32_t qpos, dpos;
};

// Passing seeds of one database chunk for queries [qb, qe) grouped by query:
// the hits of query qb + i are hits[qstart[i] .. qstart[i+1]) in no
// particular order.  The arrays belong to the device stage and stay valid
// until its second SeedHits call after this one.  A compact view carries
// only the positions (hits == nullptr): positions[k] is hit k's, and
// fetch(n, begin, count, &records) brings the full records of the n hit
// ranges [begin[i], begin[i] + count[i]) to the host, one range after the
// other, and returns (0 or a HIP error code) once they are all there, at
// *records, which stays valid until the stage's next fetch; for the
// queries whose order needs them.
struct SeedHitView {
  const SeedHit *hits = nullptr;
  const HitPosition *positions = nullptr;
  std::function<int(size_t n, const size_t *begin, const size_t *count, const SeedHit **records)>
      fetch;
  size_t count = 0;
  const uint32_t *qstart = nullptr;  // number of queries + 1
};

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
  double schedule_kernel_ms = 0;  // ordering of the gapped jobs on the device
  double traceback_kernel_ms = 0; // result build on the device
  uint64_t traceback_host_jobs = 0;  // results the device left to the host
  uint64_t gated_jobs = 0;        // learned schedule: jobs the cost model predicted
  double seed_kernel_ms[2] = {0, 0};  // seed kernels: representatives, values
  uint64_t seed_launches[2] = {0, 0};
  uint64_t gapped_launches = 0;
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
  // batched form: query blocks pipelined through chain filter (host), gapped
  // extension (device) and AddResults (host)
  bool batched = false;
  uint32_t pipeline_blocks = 8;
};

// Order keys of gapped jobs are below this bound (larger key = more work).
constexpr uint32_t kOrderKeyLimit = 4096;

// Query-length buckets of the batched form's gapped stage (residues of the
// framed query: <= 128, <= 512, <= 2048, longer).
constexpr int kLengthBuckets = 4;
inline int LengthBucket(uint32_t framed_length) {
  return framed_length <= 130 ? 0 : framed_length <= 514 ? 1 : framed_length <= 2050 ? 2 : 3;
}

// One gapped extension job as traced for the learned scheduler.
struct GappedTrace {
  uint32_t qpos;   // concatenated seed position
  uint32_t dpos;
  uint32_t chunk;
  uint8_t reverse;
  uint32_t cells;
};

// Device stages of one query chunk.  Every call returns 0 on success.
// SeedHits, GappedExtend (with JobBuffer) and Tracebacks may run on three
// host threads at once; each of the three is called from one thread at a
// time.
class PresearchDevice {
 public:
  virtual ~PresearchDevice() {}
  // Makes the query chunk resident: concatenated codes (shared delimiters;
  // coff = each query's offset), per-query ungapped cutoffs / gapped
  // triggers, and the chunk's hash lists, which the device builds itself
  // (it adds their count to st->lists and the time to
  // st->seconds_query_lists).
  virtual int BeginQueryChunk(const QuerySet &q, const std::vector<uint8_t> &concat,
                              const std::vector<uint64_t> &coff, const std::vector<int> &cutoffs,
                              const std::vector<int> &triggers, int presearch_cutoff,
                              AlignStats *st) = 0;
  // Passing seeds of queries [qb, qe) of the chunk against database chunk
  // `chunk`, grouped by query; qb is a multiple of 64 and so is qe unless it
  // is the number of queries.
  virtual int SeedHits(size_t chunk, uint32_t qb, uint32_t qe, SeedHitView *hits,
                       AlignStats *st) = 0;
  // One-sided gapped X-drop extensions of database chunk `chunk`: the first
  // n_reverse jobs extend backwards, the next n_forward forwards; qpos/dpos
  // (concatenated query / chunk positions) are replaced by the extension
  // ends and scores written; cells (optional) receives the DP cells of each.
  // The jobs are independent; the device may run them in any order:
  // keys (optional, < kOrderKeyLimit) ask for descending-key order, learned
  // asks the device to predict the keys with its cost model.
  // bucket_off (optional, kLengthBuckets + 1 entries): the jobs of each side
  // are laid out by query-length bucket, bucket b holding side-relative jobs
  // [bucket_off[b], bucket_off[b + 1]) (n_reverse == n_forward then).
  virtual int GappedExtend(size_t chunk, uint32_t n_reverse, uint32_t n_forward,
                           uint32_t *qpos, uint32_t *dpos, int *scores, uint32_t *cells,
                           const uint16_t *keys, bool learned, AlignStats *st,
                           const uint32_t *bucket_off = nullptr) = 0;
  // GappedExtend in two halves for pipelined callers: GappedPrepare hands
  // the jobs of pipeline slot `slot` (0 or 1) to the device, which may start
  // ordering them; GappedRun extends them (same arguments) and writes the
  // results back.  Prepare(slot) may overlap Run of the other slot, and the
  // two slots' Runs may be called at once from two threads (GappedExtend
  // must allow that); the host arrays stay untouched until Run(slot)
  // returns.  The defaults run the whole extension in GappedRun.
  // SetConcurrentExtensions(true) announces that the two slots' Runs will
  // overlap (a device may then give each slot its own stream).
  virtual void SetConcurrentExtensions(bool two) { (void)two; }
  virtual int GappedPrepare(int slot, size_t chunk, uint32_t n_reverse, uint32_t n_forward,
                            const uint32_t *qpos, const uint32_t *dpos, const uint16_t *keys,
                            bool learned, const uint32_t *bucket_off) {
    (void)chunk;
    (void)n_reverse;
    (void)n_forward;
    (void)qpos;
    (void)dpos;
    pending_keys_[slot & 1] = keys;
    pending_learned_[slot & 1] = learned;
    (void)bucket_off;
    return 0;
  }
  virtual int GappedRun(int slot, size_t chunk, uint32_t n_reverse, uint32_t n_forward,
                        uint32_t *qpos, uint32_t *dpos, int *scores, uint32_t *cells,
                        AlignStats *st, const uint32_t *bucket_off) {
    return GappedExtend(chunk, n_reverse, n_forward, qpos, dpos, scores, cells,
                        pending_keys_[slot & 1], pending_learned_[slot & 1], st, bucket_off);
  }
  // Result build of `jobs` of database chunk `chunk` with the result cutoff.
  // Returns kTracebackUnsupported when the device does not build results.
  virtual int Tracebacks(size_t chunk, const std::vector<TracebackJob> &jobs, int cutoff,
                         std::vector<TracebackOut> *out, AlignStats *st) {
    (void)chunk;
    (void)jobs;
    (void)cutoff;
    (void)out;
    (void)st;
    return kTracebackUnsupported;
  }
  virtual void EndQueryChunk() = 0;
  // Host memory for the job arrays of pipeline slot `slot` (at least `bytes`,
  // 16-byte aligned, valid until the next call for the same slot); a device
  // may return pinned memory so that its copies run at full speed.
  virtual uint8_t *JobBuffer(size_t slot, size_t bytes) {
    if (job_buffers_.size() <= slot) job_buffers_.resize(slot + 1);
    std::vector<uint64_t> &b = job_buffers_[slot];
    if (b.size() * 8 < bytes) b.resize((bytes + 7) / 8);
    return reinterpret_cast<uint8_t *>(b.data());
  }
  static constexpr int kTracebackUnsupported = -1;

 private:
  std::vector<std::vector<uint64_t> > job_buffers_;
  const uint16_t *pending_keys_[2] = {nullptr, nullptr};
  bool pending_learned_[2] = {false, false};
};

// Aligns all queries (split into chunks by queries_chunk_size) against all
// database chunks and appends the tabular output.  traces (optional) gets
// the gapped extensions with their DP cells when options.record_cells is
// set; concat (optional) keeps the last chunk's concatenated codes.
int AlignQueries(const QuerySet &queries, const Database &db, const AlignerOptions &options,
                 PresearchDevice *device, std::string *output, AlignStats *st,
                 std::vector<GappedTrace> *traces, std::vector<uint8_t> *concat);

}  // namespace gz
