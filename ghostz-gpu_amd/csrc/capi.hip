// capi.hip -- C ABI of the pipeline objects: queries, database, aligner.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <chrono>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <unordered_map>
#include <mutex>
#include <thread>
#include <vector>

#include "aligner.h"
#include "alphabet.h"
#include "capi_types.h"
#include "cost_model_gpu.h"
#include "database.h"
#include "gz_common.h"
#include "database_io.h"
#include "gz_runtime.h"
#include "gz_scan.h"
#include "queries.h"
#include "query_lists.h"
#include "query_lists_device.h"
#include "seg.h"
#include "statistics.h"
#include "traceback.h"

using namespace gz;

std::recursive_mutex &gz::CaptureLock() {
  static std::recursive_mutex m;
  return m;
}

struct gz_aligner_s {
  std::vector<int> matrix;
  AlignerOptions opt;
  std::unique_ptr<Aligner> aligner;
  std::unique_ptr<GpuCostModel> model;
};

namespace {

std::vector<FastaRecord> Records(const char *const *names, const char *residues,
                                 const uint64_t *offsets, uint64_t n) {
  std::vector<FastaRecord> r(n);
  for (uint64_t i = 0; i < n; ++i) {
    r[i].name = names ? names[i] : std::to_string(i);
    r[i].sequence.assign(residues + offsets[i], residues + offsets[i + 1]);
    for (char &c : r[i].sequence) c = (char)toupper((unsigned char)c);
  }
  return r;
}

bool ReadAll(const char *path, std::vector<FastaRecord> *out) {
  std::ifstream in(path);
  if (!in) return false;
  FastaRecord rec;
  while (ReadFasta(in, &rec)) out->push_back(rec);
  return true;
}

std::vector<int> Blosum() {
  std::vector<int> m;
  ParseScoreMatrix(Blosum62Text(), m);
  return m;
}

void FinishQueries(gz_queries_s *h) {
  h->name_ptrs.clear();
  for (auto &s : h->q.names) h->name_ptrs.push_back(s.c_str());
}

void FinishDatabase(gz_database_s *h) {
  h->name_ptrs.clear();
  for (auto &c : h->db.chunks) {
    h->name_ptrs.emplace_back();
    for (auto &s : c.names) h->name_ptrs.back().push_back(s.c_str());
  }
}

}  // namespace

// Output texts handed to the caller by gz_aligner_align, by data pointer,
// until gz_free.
static std::mutex g_texts_mu;
static std::unordered_map<void *, std::string *> g_texts;

extern "C" {

// ---- queries ----
gz_status gz_queries_create(const char *const *names, const char *residues,
                            const uint64_t *offsets, uint64_t n,
                            const gz_query_options *o, gz_queries_t *out) {
  GZ_REQUIRE(out && (n == 0 || (residues && offsets)), "gz_queries_create: null argument");
  QueryOptions opt;
  if (o) {
    opt.translated = o->translated != 0;
    opt.filter = o->filter != 0;
    opt.threads = o->threads;
  }
  std::vector<int> m = Blosum();
  gz_queries_s *h = new gz_queries_s();
  BuildQueries(Records(names, residues, offsets, n), opt, m.data(), kNumberLetters, &h->q);
  FinishQueries(h);
  *out = h;
  return GZ_OK;
}

gz_status gz_queries_read_fasta(const char *path, const gz_query_options *o,
                                gz_queries_t *out) {
  GZ_REQUIRE(path && out, "gz_queries_read_fasta: null argument");
  std::vector<FastaRecord> recs;
  if (!ReadAll(path, &recs)) {
    SetError(std::string("cannot read ") + path);
    return GZ_ERROR_IO;
  }
  QueryOptions opt;
  if (o) {
    opt.translated = o->translated != 0;
    opt.filter = o->filter != 0;
    opt.threads = o->threads;
  }
  std::vector<int> m = Blosum();
  gz_queries_s *h = new gz_queries_s();
  BuildQueries(recs, opt, m.data(), kNumberLetters, &h->q);
  FinishQueries(h);
  *out = h;
  return GZ_OK;
}

gz_status gz_queries_destroy(gz_queries_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_queries_view(gz_queries_t h, gz_queries_view_t *v) {
  GZ_REQUIRE(h && v, "gz_queries_view: null argument");
  v->number_queries = h->q.size();
  v->codes = h->q.codes.data();
  v->offsets = h->q.offsets.data();
  v->names = h->name_ptrs.data();
  v->lambda = h->q.lambda.data();
  v->K = h->q.K.data();
  v->translated = h->q.translated ? 1 : 0;
  v->dna_lengths = h->q.translated ? h->q.dna_lengths.data() : nullptr;
  return GZ_OK;
}

gz_status gz_queries_cutoffs(gz_queries_t h, const gz_aligner_parameters *p,
                             int *ungapped_cutoffs, int *gapped_triggers,
                             int *gapped_cutoff) {
  GZ_REQUIRE(h && (h->q.size() == 0 || (ungapped_cutoffs && gapped_triggers)),
             "gz_queries_cutoffs: null argument");
  gz_aligner_parameters d;
  gz_aligner_default_parameters(&d);
  if (!p) p = &d;
  for (size_t k = 0; k < h->q.size(); ++k) {
    KarlinParameters kp = {h->q.lambda[k], h->q.K[k], 0};
    ungapped_cutoffs[k] =
        Statistics::NormalizedCutoff2NominalCutoff(p->normalized_ungapped_cutoff, kp);
    gapped_triggers[k] = Statistics::Normalized2Nominal(p->normalized_gapped_trigger, kp);
  }
  if (gapped_cutoff)
    *gapped_cutoff = Statistics::Normalized2Nominal(p->normalized_presearch_cutoff,
                                                    Statistics::GappedBlosum62());
  return GZ_OK;
}

// ---- query hash lists ----
gz_status gz_query_lists_build(gz_queries_t h, gz_database_t db, int on_device,
                               gz_query_lists *out) {
  GZ_REQUIRE(h && db && out, "gz_query_lists_build: null argument");
  memset(out, 0, sizeof(*out));
  std::vector<uint64_t> coff;
  std::vector<uint8_t> concat;
  ConcatenatedOffsets(h->q, &coff, &concat);
  std::vector<uint32_t> hash, off, pos, qid;
  std::vector<uint8_t> check;
  if (!on_device) {
    const auto t0 = std::chrono::steady_clock::now();
    QueryLists l;
    BuildQueryLists(h->q, coff, db->db, 0, &l);
    out->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    hash.swap(l.hash);
    off.swap(l.entry_off);
    pos.swap(l.entry_pos);
    qid.swap(l.entry_qid);
    check.swap(l.entry_check);
  } else {
    hipStream_t s = nullptr;
    uint8_t *d_codes = nullptr;
    uint64_t *d_packed = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    DeviceListBuilder builder;
    DeviceQueryLists dl;
    uint64_t entries = 0;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = gz::Malloc((void **)&d_codes, std::max<size_t>(concat.size(), 16));
    if (e == hipSuccess) e = gz::Malloc((void **)&d_packed, PackedWords(concat.size()) * 8);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_codes, concat.data(), concat.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      gz_stream_s wrap{s};
      if (gz_pack_sequence_async(d_codes, concat.size(), d_packed, &wrap) != GZ_OK)
        e = hipErrorUnknown;
    }
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], s);
    if (e == hipSuccess)
      e = builder.Build(d_packed, concat.size(), h->q, coff, db->db.hash,
                        db->db.subsequence_length, s, &dl, &entries);
    if (e == hipSuccess) e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) {
      hash.resize(dl.nlists);
      off.resize((size_t)dl.nlists + 1, 0);
      pos.resize(entries);
      qid.resize(entries);
      check.resize(entries);
      if (dl.nlists) {
        e = hipMemcpyAsync(hash.data(), dl.hash, hash.size() * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(off.data(), dl.entry_off, off.size() * 4, hipMemcpyDeviceToHost, s);
      }
      if (e == hipSuccess && entries) {
        e = hipMemcpyAsync(pos.data(), dl.entry_pos, entries * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(qid.data(), dl.entry_qid, entries * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(check.data(), dl.entry_check, entries, hipMemcpyDeviceToHost, s);
      }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    out->seconds = ms / 1e3;
    for (hipEvent_t x : ev)
      if (x) (void)hipEventDestroy(x);
    if (d_codes) (void)gz::Free(d_codes);
    if (d_packed) (void)gz::Free(d_packed);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
      SetError(std::string("gz_query_lists_build: ") + hipGetErrorString(e));
      return GZ_ERROR_HIP;
    }
  }
  auto give = [](const void *src, size_t bytes) {
    void *p = malloc(std::max<size_t>(bytes, 1));
    if (p && bytes) memcpy(p, src, bytes);
    return p;
  };
  out->number_lists = hash.size();
  out->number_entries = pos.size();
  out->hash = (uint32_t *)give(hash.data(), hash.size() * 4);
  out->entry_offsets = (uint32_t *)give(off.data(), off.size() * 4);
  out->entry_positions = (uint32_t *)give(pos.data(), pos.size() * 4);
  out->entry_queries = (uint32_t *)give(qid.data(), qid.size() * 4);
  out->entry_checks = (uint8_t *)give(check.data(), check.size());
  return GZ_OK;
}

void gz_query_lists_free(gz_query_lists *l) {
  if (!l) return;
  free(l->hash);
  free(l->entry_offsets);
  free(l->entry_positions);
  free(l->entry_queries);
  free(l->entry_checks);
  memset(l, 0, sizeof(*l));
}

// ---- database ----
static DatabaseOptions DbOptions(const gz_database_options *o) {
  DatabaseOptions opt;
  if (o) {
    opt.clustering = o->clustering != 0;
    if (o->subsequence_length) opt.subsequence_length = o->subsequence_length;
    if (o->seed_threshold) opt.seed_threshold = o->seed_threshold;
    if (o->chunk_size) opt.chunk_size = o->chunk_size;
    opt.threads = o->threads;
  }
  return opt;
}

gz_status gz_database_create(const char *const *names, const char *residues,
                             const uint64_t *offsets, uint64_t n,
                             const gz_database_options *o, gz_database_t *out) {
  GZ_REQUIRE(out && (n == 0 || (residues && offsets)), "gz_database_create: null argument");
  std::vector<int> m = Blosum();
  gz_database_s *h = new gz_database_s();
  BuildDatabase(Records(names, residues, offsets, n), DbOptions(o), m.data(), kNumberLetters,
                &h->db);
  FinishDatabase(h);
  *out = h;
  return GZ_OK;
}

gz_status gz_database_read_fasta(const char *path, const gz_database_options *o,
                                 gz_database_t *out) {
  GZ_REQUIRE(path && out, "gz_database_read_fasta: null argument");
  std::vector<FastaRecord> recs;
  if (!ReadAll(path, &recs)) {
    SetError(std::string("cannot read ") + path);
    return GZ_ERROR_IO;
  }
  std::vector<int> m = Blosum();
  gz_database_s *h = new gz_database_s();
  BuildDatabase(recs, DbOptions(o), m.data(), kNumberLetters, &h->db);
  FinishDatabase(h);
  *out = h;
  return GZ_OK;
}

gz_status gz_database_load(const char *prefix, gz_database_t *out) {
  GZ_REQUIRE(prefix && out, "gz_database_load: null argument");
  std::unique_ptr<gz_database_s> h(new gz_database_s());
  std::string err;
  if (!LoadDatabase(prefix, &h->db, &err)) {
    SetError("gz_database_load: " + err);
    return GZ_ERROR_IO;
  }
  FinishDatabase(h.get());
  *out = h.release();
  return GZ_OK;
}

gz_status gz_database_save(gz_database_t h, const char *prefix) {
  GZ_REQUIRE(h && prefix, "gz_database_save: null argument");
  std::string err;
  if (!SaveDatabase(h->db, prefix, &err)) {
    SetError("gz_database_save: " + err);
    return GZ_ERROR_IO;
  }
  return GZ_OK;
}

gz_status gz_database_destroy(gz_database_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_database_info(gz_database_t h, gz_database_info_t *info) {
  GZ_REQUIRE(h && info, "gz_database_info: null argument");
  const Database &d = h->db;
  info->number_chunks = (uint32_t)d.chunks.size();
  info->total_length = d.total_length;
  info->total_sequences = d.total_sequences;
  info->max_length = d.max_length;
  info->subsequence_length = d.subsequence_length;
  memcpy(info->reduced_code_map, d.reduced_map, 25);
  memcpy(info->hash_code_map, d.hash.code_map, 25);
  info->hash_max_code = d.hash.max_code;
  info->hash_score_threshold = d.hash.threshold;
  info->hash_max_length = d.hash.max_length;
  info->hash_shift = d.hash.shift;
  for (size_t i = 0; i < 16; ++i)
    info->hash_code_scores[i] = i < d.hash.scores.size() ? d.hash.scores[i] : 0;
  info->max_mismatches = d.max_mismatches;
  return GZ_OK;
}

gz_status gz_database_chunk(gz_database_t h, uint32_t i, gz_database_chunk_view_t *v) {
  GZ_REQUIRE(h && v && i < h->db.chunks.size(), "gz_database_chunk: bad argument");
  const DbChunk &c = h->db.chunks[i];
  v->sequence = c.seq.data();
  v->sequence_length = (uint32_t)c.seq.size();
  v->number_sequences = c.NumberSequences();
  v->offsets = c.offsets.data();
  v->names = h->name_ptrs[i].data();
  v->number_hashes = (uint32_t)c.hashes.size();
  v->hashes = c.hashes.data();
  v->rep_offsets = c.rep_off.data();
  v->rep_positions = c.rep_pos.data();
  v->rep_member_offsets = c.mem_off.data();
  v->member_positions = c.mem_pos.data();
  v->value_offsets = c.val_off.data();
  v->value_positions = c.val_pos.data();
  v->number_reps = c.rep_pos.size();
  v->number_members = c.mem_pos.size();
  v->number_values = c.val_pos.size();
  v->clustering = c.clustering ? 1 : 0;
  return GZ_OK;
}

// ---- result build ----
static_assert(sizeof(gz_traceback_job) == sizeof(TracebackJob), "job layout");
static_assert(sizeof(gz_traceback_result) == sizeof(TracebackOut), "result layout");

gz_status gz_build_results(const uint64_t *pq, const uint64_t *pd, const gz_traceback_job *jobs,
                           size_t n, uint32_t max_query_length, uint32_t max_subject_length,
                           const int *d_matrix, uint32_t nl, int gap_open, int gap_ext,
                           int cutoff, gz_traceback_result *results, gz_stream_t stream) {
  GZ_REQUIRE(pq && pd && d_matrix && (n == 0 || (jobs && results)) && nl <= 32,
             "gz_build_results: bad argument");
  GZ_REQUIRE(n < (1ull << 31), "gz_build_results: too many jobs");
  if (n == 0) return GZ_OK;
  hipStream_t s = stream ? stream->s : nullptr;
  const LaneCaps caps = TracebackCaps(max_query_length, max_subject_length);
  const size_t lanes = std::min<size_t>((n + 63) / 64 * 64, (size_t)1 << 14);
  void *ws = nullptr;
  const size_t a_jobs = AlignUp(n * sizeof(TracebackJob), 256);
  const size_t a_outs = AlignUp(n * sizeof(TracebackOut), 256);
  GZ_HIP_CHECK(gz::Malloc(&ws, a_jobs + a_outs + lanes * caps.bytes_per_lane));
  hipError_t e = hipMemcpyAsync(ws, jobs, n * sizeof(TracebackJob), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = TracebackLaunch(pq, pd, (const TracebackJob *)ws, (uint32_t)n, d_matrix, nl, gap_open,
                        gap_ext, cutoff, (uint8_t *)ws + a_jobs + a_outs, caps, (uint32_t)lanes,
                        (TracebackOut *)((char *)ws + a_jobs), s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(results, (char *)ws + a_jobs, n * sizeof(TracebackOut),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)gz::Free(ws);
  GZ_HIP_CHECK(e);
  return GZ_OK;
}

gz_status gz_build_result_host(const uint8_t *query, uint32_t query_length,
                               const uint8_t *database, uint32_t hit_q, uint32_t hit_d,
                               int cutoff, gz_traceback_result *result) {
  GZ_REQUIRE(query && database && result && hit_q >= 1 && hit_q < query_length,
             "gz_build_result_host: bad argument");
  std::vector<int> m = Blosum();
  SearchParameters p;
  p.matrix = m.data();
  p.nl = kNumberLetters;
  const TracebackOut o = TracebackOneHost(query, query_length, database, hit_q, hit_d, p, cutoff);
  memcpy(result, &o, sizeof(o));
  return GZ_OK;
}

// ---- aligner ----
void gz_aligner_default_parameters(gz_aligner_parameters *p) {
  memset(p, 0, sizeof(*p));
  p->gap_open = 11;
  p->gap_extension = 1;
  p->normalized_ungapped_cutoff = 7.0f;
  p->normalized_gapped_trigger = 22.0f;
  p->normalized_presearch_cutoff = 15.0f;
  p->normalized_result_cutoff = 25.0f;
  p->max_results = 10;
  p->max_results_per_subject = 1;
  p->queries_chunk_size = 1ull << 27;
  p->threads = 0;
  p->schedule = GZ_SCHEDULE_PLAIN;
  p->record_traces = 0;
  p->batched = 0;
  p->pipeline_blocks = 8;
}

gz_status gz_aligner_create(const gz_aligner_parameters *p, gz_aligner_t *out) {
  GZ_REQUIRE(out, "gz_aligner_create: null argument");
  gz_aligner_parameters d;
  gz_aligner_default_parameters(&d);
  if (!p) p = &d;
  gz_aligner_s *h = new gz_aligner_s();
  h->matrix = Blosum();
  AlignerOptions &o = h->opt;
  o.search.matrix = h->matrix.data();
  o.search.nl = kNumberLetters;
  o.search.gap_open = p->gap_open;
  o.search.gap_extension = p->gap_extension;
  o.search.normalized_ungapped_cutoff = p->normalized_ungapped_cutoff;
  o.search.normalized_gapped_trigger = p->normalized_gapped_trigger;
  o.search.normalized_presearch_cutoff = p->normalized_presearch_cutoff;
  o.search.normalized_result_cutoff = p->normalized_result_cutoff;
  o.search.max_results = p->max_results;
  o.search.max_results_per_subject = p->max_results_per_subject;
  o.queries_chunk_size = p->queries_chunk_size ? p->queries_chunk_size : (1ull << 27);
  o.threads = p->threads;
  o.schedule = (Schedule)p->schedule;
  o.record_cells = p->record_traces != 0;
  o.batched = p->batched != 0;
  o.pipeline_blocks = p->pipeline_blocks ? p->pipeline_blocks : 8;
  h->aligner.reset(new Aligner(o));
  *out = h;
  return GZ_OK;
}

gz_status gz_aligner_destroy(gz_aligner_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_aligner_set_database(gz_aligner_t h, gz_database_t db) {
  GZ_REQUIRE(h && db, "gz_aligner_set_database: null argument");
  GZ_HIP_CHECK(h->aligner->SetDatabase(&db->db));
  return GZ_OK;
}

gz_status gz_aligner_graph_counts(gz_aligner_t h, uint64_t *instantiations,
                                  uint64_t *updates) {
  GZ_REQUIRE(h && instantiations && updates, "gz_aligner_graph_counts: null argument");
  h->aligner->GraphCounts(instantiations, updates);
  return GZ_OK;
}

gz_status gz_aligner_tuner_state(gz_aligner_t h, int *chosen, double *ms_per_job,
                                 uint32_t *trials) {
  GZ_REQUIRE(h, "gz_aligner_tuner_state: null argument");
  const GappedTuner t = h->aligner->Tuner();
  for (int b = 0; b < kLengthBuckets; ++b) {
    bool tried = false;
    for (int c = 0; c < GappedTuner::kConfigs; ++c) {
      tried = tried || t.trials[b][c] > 0;
      if (ms_per_job)
        ms_per_job[b * GappedTuner::kConfigs + c] =
            t.jobs[b][c] > 0 ? t.ms[b][c] / t.jobs[b][c] : 0.0;
      if (trials) trials[b * GappedTuner::kConfigs + c] = (uint32_t)t.trials[b][c];
    }
    if (chosen) chosen[b] = tried ? t.Pick(b) : -1;
  }
  return GZ_OK;
}

gz_status gz_aligner_set_cost_model(gz_aligner_t h, const float *weights, size_t count) {
  GZ_REQUIRE(h && weights, "gz_aligner_set_cost_model: null argument");
  CostModelWeights w;
  if (!UnflattenWeights(weights, count, &w)) {
    SetError("gz_aligner_set_cost_model: wrong weight count");
    return GZ_ERROR_INVALID_ARGUMENT;
  }
  std::unique_ptr<GpuCostModel> model(new GpuCostModel(w));
  GZ_HIP_CHECK(model->status());
  h->aligner->SetCostModel(model.get());
  h->opt.cost_model = model.get();
  h->model = std::move(model);
  return GZ_OK;
}

gz_status gz_aligner_align(gz_aligner_t h, gz_queries_t q, char **output,
                           size_t *output_length, gz_align_stats *stats) {
  GZ_REQUIRE(h && q && output && output_length, "gz_aligner_align: null argument");
  GZ_REQUIRE(h->aligner->HasDatabase(), "gz_aligner_align: no database set");
  std::unique_ptr<std::string> out(new std::string);
  AlignStats st;
  GZ_HIP_CHECK(h->aligner->Align(q->q, out.get(), &st));
  // the text is handed over without a copy; gz_free releases it
  *output_length = out->size();
  *output = &(*out)[0];
  {
    std::lock_guard<std::mutex> lock(g_texts_mu);
    g_texts[*output] = out.release();
  }
  if (stats) {
    stats->seconds_total = st.seconds_total;
    stats->seconds_query_lists = st.seconds_query_lists;
    stats->seconds_upload = st.seconds_upload;
    stats->seconds_seed = st.seconds_seed;
    stats->seconds_chain = st.seconds_chain;
    stats->seconds_gapped = st.seconds_gapped;
    stats->seconds_add = st.seconds_add;
    stats->seconds_build = st.seconds_build;
    stats->seconds_output = st.seconds_output;
    stats->query_chunks = st.query_chunks;
    stats->lists = st.lists;
    stats->units_rep = st.units_rep;
    stats->units_value = st.units_value;
    stats->seed_hits = st.seed_hits;
    stats->chained_hits = st.chained_hits;
    stats->gapped_extensions = st.gapped_extensions;
    stats->gapped_cells = st.gapped_cells;
    stats->results = st.results;
    stats->sequential_build_chunks = st.sequential_build_chunks;
    stats->gapped_kernel_ms = st.gapped_kernel_ms;
    stats->seed_rep_kernel_ms = st.seed_kernel_ms[0];
    stats->seed_value_kernel_ms = st.seed_kernel_ms[1];
    stats->seed_rep_launches = st.seed_launches[0];
    stats->seed_value_launches = st.seed_launches[1];
    stats->gapped_launches = st.gapped_launches;
    stats->schedule_kernel_ms = st.schedule_kernel_ms;
    stats->traceback_kernel_ms = st.traceback_kernel_ms;
    stats->traceback_host_jobs = st.traceback_host_jobs;
    stats->gated_jobs = st.gated_jobs;
  }
  return GZ_OK;
}

uint64_t gz_aligner_traces(gz_aligner_t h, gz_gapped_trace *out, uint64_t capacity) {
  if (!h) return 0;
  const std::vector<GappedTrace> &t = h->aligner->Traces();
  if (out) {
    for (uint64_t i = 0; i < t.size() && i < capacity; ++i) {
      out[i].query_position = t[i].qpos;
      out[i].database_position = t[i].dpos;
      out[i].chunk = t[i].chunk;
      out[i].reverse = t[i].reverse;
      out[i].cells = t[i].cells;
    }
  }
  return t.size();
}

gz_status gz_aligner_last_concatenated_queries(gz_aligner_t h, const uint8_t **codes,
                                               uint64_t *length) {
  GZ_REQUIRE(h && codes && length, "null argument");
  *codes = h->aligner->LastConcatenatedQueries().data();
  *length = h->aligner->LastConcatenatedQueries().size();
  return GZ_OK;
}

namespace {

// Aligns a query FASTA against a database (either source) into a file.
gz_status AlignInto(const char *queries_path, const char *output_path, int translated,
                    const gz_aligner_parameters *p, gz_database_t db, gz_status st) {
  gz_query_options qo = {translated, 1, 0};
  gz_queries_t q = nullptr;
  gz_aligner_t a = nullptr;
  if (st == GZ_OK) st = gz_queries_read_fasta(queries_path, &qo, &q);
  if (st == GZ_OK) st = gz_aligner_create(p, &a);
  if (st == GZ_OK) st = gz_aligner_set_database(a, db);
  char *out = nullptr;
  size_t len = 0;
  if (st == GZ_OK) st = gz_aligner_align(a, q, &out, &len, nullptr);
  if (st == GZ_OK) {
    std::ofstream os(output_path, std::ios::binary);
    if (!os) {
      SetError(std::string("cannot write ") + output_path);
      st = GZ_ERROR_IO;
    } else {
      os.write(out, (std::streamsize)len);
    }
  }
  free(out);
  gz_aligner_destroy(a);
  gz_queries_destroy(q);
  gz_database_destroy(db);
  return st;
}

}  // namespace

gz_status gz_align_fasta(const char *queries_path, const char *database_path,
                         const char *output_path, int translated,
                         const gz_aligner_parameters *p,
                         const gz_database_options *dbo) {
  GZ_REQUIRE(queries_path && database_path && output_path, "gz_align_fasta: null argument");
  gz_database_t db = nullptr;
  const gz_status st = gz_database_read_fasta(database_path, dbo, &db);
  return AlignInto(queries_path, output_path, translated, p, db, st);
}

gz_status gz_align_main(const char *queries_path, const char *database_prefix,
                        const char *output_path, int translated,
                        const gz_aligner_parameters *p) {
  GZ_REQUIRE(queries_path && database_prefix && output_path, "gz_align_main: null argument");
  // AlignMain::Run (src/align_main.cpp:46-62): both inputs must exist
  std::ifstream qi(queries_path);
  if (!qi) {
    SetError("Query file does not exist.");
    return GZ_ERROR_IO;
  }
  std::ifstream di((std::string(database_prefix) + ".inf").c_str());
  if (!di) {
    SetError("Database file does not exist.");
    return GZ_ERROR_IO;
  }
  gz_database_t db = nullptr;
  const gz_status st = gz_database_load(database_prefix, &db);
  return AlignInto(queries_path, output_path, translated, p, db, st);
}

void gz_free(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lock(g_texts_mu);
    auto it = g_texts.find(p);
    if (it != g_texts.end()) {
      delete it->second;
      g_texts.erase(it);
      return;
    }
  }
  free(p);
}

// ---- host-side preparation pieces exposed for tests ----
int gz_seg_mask(const char *seq, char *out) {
  static thread_local Seg seg(12);
  std::string masked;
  const int r = seg.Mask(seq, &masked);
  if (r == 0) memcpy(out, masked.data(), masked.size());
  return r;
}

double gz_seg_ln_factorial(int n) { return SegLnFactorial(n); }

int gz_translate(const char *dna, char *out, uint32_t *lengths) {
  std::vector<std::string> f;
  TranslateSixFrames(dna, f);
  size_t o = 0;
  for (size_t i = 0; i < 6; ++i) {
    if (i < f.size()) {
      memcpy(out + o, f[i].data(), f[i].size());
      lengths[i] = (uint32_t)f[i].size();
      o += f[i].size();
    } else {
      lengths[i] = 0;
    }
  }
  return (int)f.size();
}

void gz_karlin_block(const double *p_at_min, int min, int max, float *lambda, float *K,
                     float *H) {
  KarlinBlockCalc(p_at_min, min, max, lambda, K, H);
}

int gz_query_karlin(const uint8_t *seq, uint32_t len, float *lambda, float *K, float *H) {
  std::vector<int> m = Blosum();
  Statistics s;
  KarlinParameters p;
  const bool ok = s.QueryUngapped(seq, len, m.data(), kNumberLetters, &p);
  *lambda = p.lambda;
  *K = p.K;
  *H = p.H;
  return ok ? 1 : 0;
}

uint64_t gz_search_space(uint64_t qlen, uint64_t dblen, uint64_t dbn) {
  return Statistics::SearchSpace(qlen, dblen, dbn, Statistics::GappedBlosum62(), 1.9, -30);
}

}  // extern "C"
