// aligner.hip -- see aligner.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <numeric>
#include <sstream>
#include <thread>

#include "aligner.h"
#include "cost_model.h"
#include "gz_runtime.h"
#include "gz_common.h"
#include "query_lists.h"

namespace gz {

namespace {

double Now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <class F>
void ParallelBlocks(size_t n, int threads, size_t grain, F f) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  threads = std::max(1, std::min<int>(threads, (int)((n + grain - 1) / std::max<size_t>(grain, 1))));
  std::atomic<size_t> next(0);
  auto work = [&]() {
    for (;;) {
      const size_t b = next.fetch_add(grain);
      if (b >= n) break;
      f(b, std::min(n, b + grain));
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
}

template <class T>
hipError_t Upload(T **d, const std::vector<T> &h, hipStream_t s) {
  hipError_t e = hipMalloc((void **)d, std::max<size_t>(h.size() * sizeof(T), 16));
  if (e != hipSuccess) return e;
  if (!h.empty()) e = hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
  return e;
}

// packs host codes through a temporary device copy
hipError_t UploadPacked(uint64_t **d, const std::vector<uint8_t> &codes, hipStream_t s) {
  uint8_t *tmp = nullptr;
  hipError_t e = hipMalloc((void **)&tmp, std::max<size_t>(codes.size(), 16));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(tmp, codes.data(), codes.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMalloc((void **)d, PackedWords(codes.size()) * 8);
  if (e == hipSuccess) {
    gz_stream_s wrap{s};
    if (gz_pack_sequence_async(tmp, codes.size(), *d, &wrap) != GZ_OK) e = hipErrorUnknown;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  return e;
}

struct KeyOrder {
  const SeedHit *h;
  bool operator()(uint32_t a, uint32_t b) const {
    const SeedHit &x = h[a], &y = h[b];
    if (x.list != y.list) return x.list < y.list;
    if (x.rv != y.rv) return x.rv < y.rv;
    if (x.member != y.member) return x.member < y.member;
    return x.entry < y.entry;
  }
};

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
      if (p) (void)hipFree(p);
  }
};

Aligner::Aligner(const AlignerOptions &opt) : opt_(opt) {}

Aligner::~Aligner() {
  chunks_.clear();
  if (gapped_) gz_gapped_extender_destroy(gapped_);
  if (d_matrix_) (void)hipFree(d_matrix_);
  if (d_reduced_) (void)hipFree(d_reduced_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

hipError_t Aligner::SetDatabase(const Database *db) {
  db_ = db;
  if (!stream_) GZ_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
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
  if (d_reduced_) GZ_TRY(hipFree(d_reduced_));
  GZ_TRY(Upload(&d_reduced_, rm, stream_));
  GZ_TRY(hipStreamSynchronize(stream_));
  if (!gapped_ && gz_gapped_extender_create(&gapped_) != GZ_OK) return hipErrorUnknown;
  return hipSuccess;
}

hipError_t Aligner::Align(const QuerySet &queries, std::string *output, AlignStats *st) {
  // query chunks (Queries::SetQueries, src/queries.cpp:183-224)
  const size_t n = queries.size();
  size_t i = 0;
  traces_.clear();
  const double t0 = Now();
  while (i < n) {
    uint64_t sum = 0;
    size_t j = i;
    while (j < n) {
      const uint64_t next = sum + queries.Length(j);
      if (next >= opt_.queries_chunk_size) break;
      sum = next;
      ++j;
    }
    if (j == i) j = i + 1;
    QuerySet part;
    part.translated = queries.translated;
    part.max_length = 0;
    part.offsets.push_back(0);
    for (size_t k = i; k < j; ++k) {
      part.names.push_back(queries.names[k]);
      part.codes.insert(part.codes.end(), queries.Sequence(k), queries.Sequence(k) + queries.Length(k));
      part.offsets.push_back(part.codes.size());
      part.lambda.push_back(queries.lambda[k]);
      part.K.push_back(queries.K[k]);
      if (queries.translated) part.dna_lengths.push_back(queries.dna_lengths[k]);
      part.max_length = std::max(part.max_length, queries.Length(k));
    }
    GZ_TRY(AlignChunk(part, output, st));
    if (st) ++st->query_chunks;
    i = j;
  }
  if (st) st->seconds_total = Now() - t0;
  return hipSuccess;
}

hipError_t Aligner::AlignChunk(const QuerySet &q, std::string *output, AlignStats *st) {
  AlignStats dummy;
  if (!st) st = &dummy;
  const SearchParameters &P = opt_.search;
  const size_t nq = q.size();
  const int threads = opt_.threads > 0 ? opt_.threads : (int)std::thread::hardware_concurrency();
  hipStream_t s = stream_;
  // cutoffs (AlignerGpu::SetQueriesData, src/aligner_gpu.cpp:213-242)
  const KarlinParameters gk = Statistics::GappedBlosum62();
  const int presearch_cutoff = Statistics::Normalized2Nominal(P.normalized_presearch_cutoff, gk);
  const int result_cutoff = Statistics::Normalized2Nominal(P.normalized_result_cutoff, gk);
  std::vector<int> cut(nq), trig(nq);
  for (size_t k = 0; k < nq; ++k) {
    KarlinParameters kp = {q.lambda[k], q.K[k], 0};
    cut[k] = Statistics::NormalizedCutoff2NominalCutoff(P.normalized_ungapped_cutoff, kp);
    trig[k] = Statistics::Normalized2Nominal(P.normalized_gapped_trigger, kp);
  }
  double t = Now();
  std::vector<uint64_t> coff;
  ConcatenatedOffsets(q, &coff, &concat_);
  QueryLists lists;
  BuildQueryLists(q, coff, *db_, threads, &lists);
  st->seconds_query_lists += Now() - t;
  st->lists += lists.NumberLists();

  t = Now();
  uint64_t *d_q = nullptr;
  int *d_cut = nullptr, *d_trig = nullptr;
  uint32_t *d_lhash = nullptr, *d_eoff = nullptr, *d_epos = nullptr, *d_eqid = nullptr;
  uint8_t *d_echeck = nullptr;
  GZ_TRY(UploadPacked(&d_q, concat_, s));
  GZ_TRY(Upload(&d_cut, cut, s));
  GZ_TRY(Upload(&d_trig, trig, s));
  GZ_TRY(Upload(&d_lhash, lists.hash, s));
  GZ_TRY(Upload(&d_eoff, lists.entry_off, s));
  GZ_TRY(Upload(&d_epos, lists.entry_pos, s));
  GZ_TRY(Upload(&d_eqid, lists.entry_qid, s));
  GZ_TRY(Upload(&d_echeck, lists.entry_check, s));
  GZ_TRY(hipStreamSynchronize(s));
  st->seconds_upload += Now() - t;

  DeviceQueryLists dq;
  dq.packed = d_q;
  dq.cutoffs = d_cut;
  dq.triggers = d_trig;
  dq.nlists = (uint32_t)lists.NumberLists();
  dq.hash = d_lhash;
  dq.entry_off = d_eoff;
  dq.entry_pos = d_epos;
  dq.entry_qid = d_eqid;
  dq.entry_check = d_echeck;
  SeedStageParams sp;
  sp.matrix = d_matrix_;
  sp.nl = P.nl;
  sp.delimiter = kDelimiter;
  sp.reduced_map = d_reduced_;
  sp.subsequence_length = db_->subsequence_length;
  {
    const float rate = 0.8f;
    const uint32_t L = db_->subsequence_length;
    const char threshold = L - rate * L;  // SeedSearcherGpu::Reset (:58-69)
    sp.distance_threshold = threshold;
  }
  sp.max_mismatches = db_->max_mismatches;

  if (gz_gapped_extender_set_max_sequence_length(gapped_, q.max_length) != GZ_OK ||
      gz_gapped_extender_set_queries(gapped_, kDelimiter, d_q) != GZ_OK ||
      gz_gapped_extender_set_score_parameters(gapped_, d_matrix_, P.nl, P.gap_open,
                                              P.gap_extension, presearch_cutoff) != GZ_OK)
    return hipErrorUnknown;

  std::vector<std::vector<Presearched> > pre(nq);
  hipEvent_t ev0, ev1;
  GZ_TRY(hipEventCreate(&ev0));
  GZ_TRY(hipEventCreate(&ev1));
  for (size_t ci = 0; ci < db_->chunks.size(); ++ci) {
    const DbChunk &chunk = db_->chunks[ci];
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
    sp.clustering = chunk.clustering;
    // 1. seed stage
    t = Now();
    std::vector<SeedHit> hits;
    SeedStageStats ss;
    GZ_TRY(seed_.Run(dq, ddb, sp, s, &hits, &ss));
    st->units_rep += ss.units_rep;
    st->units_value += ss.units_value;
    st->seed_hits += hits.size();
    st->seconds_seed += Now() - t;

    // 2. per-query enumeration order and chain filter
    t = Now();
    std::vector<uint32_t> qstart(nq + 1, 0);
    for (const SeedHit &h : hits) ++qstart[h.qid + 1];
    for (size_t k = 0; k < nq; ++k) qstart[k + 1] += qstart[k];
    std::vector<uint32_t> byq(hits.size());
    {
      std::vector<uint32_t> fill(qstart.begin(), qstart.end() - 1);
      for (uint32_t k = 0; k < hits.size(); ++k) byq[fill[hits[k].qid]++] = k;
    }
    std::vector<std::vector<Hit> > chained(nq);
    ParallelBlocks(nq, threads, 64, [&](size_t b, size_t e) {
      std::vector<uint64_t> scratch;
      KeyOrder order{hits.data()};
      for (size_t qi = b; qi < e; ++qi) {
        const uint32_t s0 = qstart[qi], s1 = qstart[qi + 1];
        if (s0 == s1) continue;
        std::sort(byq.begin() + s0, byq.begin() + s1, order);
        std::vector<Hit> &hv = chained[qi];
        hv.resize(s1 - s0);
        for (uint32_t k = s0; k < s1; ++k) {
          const SeedHit &h = hits[byq[k]];
          hv[k - s0].q = h.qpos - (uint32_t)coff[qi];
          hv[k - s0].d = h.dpos;
        }
        ChainFilter(q.Sequence(qi), q.Length(qi), cut[qi], chunk.seq.data(), P.matrix, P.nl,
                    hv, scratch);
      }
    });
    std::vector<SeedHit>().swap(hits);
    std::vector<uint32_t> gstart(nq + 1, 0);
    for (size_t k = 0; k < nq; ++k) gstart[k + 1] = gstart[k] + (uint32_t)chained[k].size();
    const uint32_t G = gstart[nq];
    st->chained_hits += G;
    st->seconds_chain += Now() - t;

    // 3. gapped extension of both sides on the device
    t = Now();
    std::vector<uint32_t> rq(G), rd(G), fq(G), fd(G);
    for (size_t qi = 0; qi < nq; ++qi)
      for (uint32_t k = 0; k < chained[qi].size(); ++k) {
        const uint32_t g = gstart[qi] + k;
        const Hit &h = chained[qi][k];
        rq[g] = (uint32_t)coff[qi] + h.q - 1;
        rd[g] = h.d - 1;
        fq[g] = (uint32_t)coff[qi] + h.q;
        fd[g] = h.d;
      }
    std::vector<int> rs(G), fs(G);
    std::vector<uint32_t> rc, fc;
    if (G > 0) {
      if (gz_gapped_extender_set_database(gapped_, dc.packed) != GZ_OK) return hipErrorUnknown;
      // schedule: the order extensions are handed to the lanes
      std::vector<uint32_t> perm_r(G), perm_f(G);
      std::iota(perm_r.begin(), perm_r.end(), 0);
      std::iota(perm_f.begin(), perm_f.end(), 0);
      if (opt_.schedule != Schedule::kPlain) {
        std::vector<float> cr(G), cf(G);
        if (opt_.schedule == Schedule::kLearned && opt_.cost_model) {
          opt_.cost_model->Predict(concat_.data(), chunk.seq.data(), rq.data(), rd.data(), G,
                                   true, cr.data());
          opt_.cost_model->Predict(concat_.data(), chunk.seq.data(), fq.data(), fd.data(), G,
                                   false, cf.data());
        } else {
          for (size_t qi = 0; qi < nq; ++qi)
            for (uint32_t k = gstart[qi]; k < gstart[qi + 1]; ++k) {
              cr[k] = (float)q.DistanceFromStart(qi, rq[k] - (uint32_t)coff[qi]);
              cf[k] = (float)q.DistanceFromEnd(qi, fq[k] - (uint32_t)coff[qi]);
            }
        }
        std::stable_sort(perm_r.begin(), perm_r.end(),
                         [&](uint32_t a, uint32_t b) { return cr[a] > cr[b]; });
        std::stable_sort(perm_f.begin(), perm_f.end(),
                         [&](uint32_t a, uint32_t b) { return cf[a] > cf[b]; });
      }
      std::vector<uint32_t> aq(2 * G), ad(2 * G);
      for (uint32_t k = 0; k < G; ++k) {
        aq[k] = rq[perm_r[k]];
        ad[k] = rd[perm_r[k]];
        aq[G + k] = fq[perm_f[k]];
        ad[G + k] = fd[perm_f[k]];
      }
      uint32_t *d_pq = nullptr, *d_pd = nullptr, *d_cells = nullptr;
      int *d_sc = nullptr;
      GZ_TRY(hipMalloc((void **)&d_pq, 2 * G * 4));
      GZ_TRY(hipMalloc((void **)&d_pd, 2 * G * 4));
      GZ_TRY(hipMalloc((void **)&d_sc, 2 * G * 4));
      if (opt_.record_cells) GZ_TRY(hipMalloc((void **)&d_cells, 2 * G * 4));
      GZ_TRY(hipMemcpyAsync(d_pq, aq.data(), 2 * G * 4, hipMemcpyHostToDevice, s));
      GZ_TRY(hipMemcpyAsync(d_pd, ad.data(), 2 * G * 4, hipMemcpyHostToDevice, s));
      gz_stream_s wrap{s};
      GZ_TRY(hipEventRecord(ev0, s));
      if (gz_gapped_extender_extend_one_side_score_only_cells_async(
              gapped_, G, 1, d_pq, d_pd, d_sc, d_cells, &wrap) != GZ_OK ||
          gz_gapped_extender_extend_one_side_score_only_cells_async(
              gapped_, G, 0, d_pq + G, d_pd + G, d_sc + G, d_cells ? d_cells + G : nullptr,
              &wrap) != GZ_OK)
        return hipErrorUnknown;
      GZ_TRY(hipEventRecord(ev1, s));
      std::vector<int> sc(2 * G);
      GZ_TRY(hipMemcpyAsync(aq.data(), d_pq, 2 * G * 4, hipMemcpyDeviceToHost, s));
      GZ_TRY(hipMemcpyAsync(ad.data(), d_pd, 2 * G * 4, hipMemcpyDeviceToHost, s));
      GZ_TRY(hipMemcpyAsync(sc.data(), d_sc, 2 * G * 4, hipMemcpyDeviceToHost, s));
      std::vector<uint32_t> cells;
      if (d_cells) {
        cells.resize(2 * G);
        GZ_TRY(hipMemcpyAsync(cells.data(), d_cells, 2 * G * 4, hipMemcpyDeviceToHost, s));
      }
      GZ_TRY(hipStreamSynchronize(s));
      uint32_t fails = 0;
      if (gz_gapped_extender_failures(gapped_, &fails) != GZ_OK || fails != 0) return hipErrorUnknown;
      float ms = 0;
      GZ_TRY(hipEventElapsedTime(&ms, ev0, ev1));
      st->gapped_kernel_ms += ms;
      (void)hipFree(d_pq);
      (void)hipFree(d_pd);
      (void)hipFree(d_sc);
      if (d_cells) (void)hipFree(d_cells);
      for (uint32_t k = 0; k < G; ++k) {
        rq[perm_r[k]] = aq[k];
        rd[perm_r[k]] = ad[k];
        rs[perm_r[k]] = sc[k];
        fq[perm_f[k]] = aq[G + k];
        fd[perm_f[k]] = ad[G + k];
        fs[perm_f[k]] = sc[G + k];
      }
      if (!cells.empty()) {
        for (uint32_t k = 0; k < G; ++k) {
          GappedTrace tr;
          tr.chunk = (uint32_t)ci;
          tr.reverse = 1;
          tr.qpos = 0;
          tr.dpos = 0;
          tr.cells = cells[k];
          const uint32_t g = perm_r[k];
          // seed positions before extension
          const uint32_t qi = (uint32_t)(std::upper_bound(gstart.begin(), gstart.end(), g) -
                                         gstart.begin()) - 1;
          const Hit &h = chained[qi][g - gstart[qi]];
          tr.qpos = (uint32_t)coff[qi] + h.q - 1;
          tr.dpos = h.d - 1;
          traces_.push_back(tr);
          st->gapped_cells += cells[k];
        }
        for (uint32_t k = 0; k < G; ++k) {
          GappedTrace tr;
          const uint32_t g = perm_f[k];
          const uint32_t qi = (uint32_t)(std::upper_bound(gstart.begin(), gstart.end(), g) -
                                         gstart.begin()) - 1;
          const Hit &h = chained[qi][g - gstart[qi]];
          tr.chunk = (uint32_t)ci;
          tr.reverse = 0;
          tr.qpos = (uint32_t)coff[qi] + h.q;
          tr.dpos = h.d;
          tr.cells = cells[G + k];
          traces_.push_back(tr);
          st->gapped_cells += cells[G + k];
        }
      }
    }
    st->gapped_extensions += 2 * (uint64_t)G;
    st->seconds_gapped += Now() - t;

    // 4. presearched results and AddResults per query
    t = Now();
    ParallelBlocks(nq, threads, 64, [&](size_t b, size_t e) {
      std::vector<std::vector<AlignmentPosition> > starts;
      std::vector<Presearched> tmp;
      for (size_t qi = b; qi < e; ++qi) {
        tmp.clear();
        for (uint32_t g = gstart[qi]; g < gstart[qi + 1]; ++g) {
          const Hit &h = chained[qi][g - gstart[qi]];
          Presearched r;
          r.chunk = -1;
          r.subject = 0;
          r.hit.q = h.q;
          r.hit.d = h.d;
          r.start.q = rq[g] - (uint32_t)coff[qi];
          r.start.d = rd[g];
          r.end.q = fq[g] - (uint32_t)coff[qi];
          r.end.d = fd[g];
          r.score = rs[g] + fs[g];
          r.hit_count = 1;
          tmp.push_back(r);
        }
        if (!tmp.empty() || !pre[qi].empty())
          AddResults(chunk, (int)ci, P, starts, tmp, pre[qi]);
      }
    });
    st->seconds_add += Now() - t;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  void *frees[] = {d_q, d_cut, d_trig, d_lhash, d_eoff, d_epos, d_eqid, d_echeck};
  for (void *p : frees) (void)hipFree(p);

  // BuildResults and WriteOutput
  t = Now();
  std::vector<std::vector<AlignmentResult> > results;
  BuildResults(q, *db_, P, result_cutoff, pre, results, threads, &st->sequential_build_chunks);
  for (auto &r : results) st->results += r.size();
  st->seconds_build += Now() - t;
  t = Now();
  std::ostringstream os;
  WriteOutput(os, q, *db_, results);
  output->append(os.str());
  st->seconds_output += Now() - t;
  return hipSuccess;
}

}  // namespace gz
