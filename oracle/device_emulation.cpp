/*
 * device_emulation.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Runs the product's host flow (AlignQueries: query chunking, hash lists,
 * per-query seed order, chain filter, schedule, AddResults, BuildResults,
 * WriteOutput) with its two device stages emulated on the CPU by the
 * oracle's leaf functions:
 *   SeedHits     -- every (list, representative/value, member, entry) unit
 *                   the device seed kernel evaluates, in any order, with the
 *                   same distance / ungapped-trigger tests (gzo_distance,
 *                   gzo_ungapped_with_trigger);
 *   GappedExtend -- gzo_gapped_one_side per job.
 * This lets CPU tests (and sanitizer builds) exercise the product's host
 * code without a GPU.  The product library never contains this file.
 */
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <charconv>
#include <mutex>
#include <string>
#include <vector>

#include "alphabet.h"
#include "database.h"
#include "ghostz_hip.h"
#include "gz_oracle.h"
#include "presearch.h"
#include "queries.h"
#include "results.h"
#include "ungapped.h"

using namespace gz;

namespace {

// every n-th traceback (n > 0) reports a stale read: the host must then
// rebuild the chunk in the reference's order
int g_forced_stale = 0;

class CpuDevice : public PresearchDevice {
 public:
  CpuDevice(const Database &db, const SearchParameters &p) : db_(db), p_(p) {}

  int BeginQueryChunk(const QuerySet &q, const std::vector<uint8_t> &concat,
                      const std::vector<uint64_t> &coff, const std::vector<int> &cut,
                      const std::vector<int> &trig, int presearch_cutoff,
                      AlignStats *st) override {
    concat_ = &concat;
    BuildQueryLists(q, coff, db_, 0, &lists_);
    st->lists += lists_.NumberLists();
    cut_ = cut;
    trig_ = trig;
    presearch_cutoff_ = presearch_cutoff;
    return 0;
  }

  int SeedHits(size_t ci, uint32_t qb, uint32_t qe, SeedHitView *view,
               AlignStats *st) override {
    if (qb > qe || (qb & 63) != 0 || ((qe & 63) != 0 && qe != cut_.size())) return -2;
    const uint32_t nq = qe - qb;
    const int slot = flip_;
    flip_ ^= 1;
    std::vector<SeedHit> *hits = &raw_;
    const DbChunk &c = db_.chunks[ci];
    const uint8_t *qs = concat_->data();
    const uint8_t *ds = c.seq.data();
    const uint32_t L = db_.subsequence_length;
    const float rate = 0.8f;
    const char threshold = L - rate * L;
    const QueryLists &ql = lists_;
    hits->clear();
    for (uint32_t l = 0; l < ql.NumberLists(); ++l) {
      const uint32_t h = ql.hash[l];
      const auto it = std::lower_bound(c.hashes.begin(), c.hashes.end(), h);
      if (it == c.hashes.end() || *it != h) continue;
      const uint32_t hi = (uint32_t)(it - c.hashes.begin());
      const uint32_t eb = ql.entry_off[l], ee = ql.entry_off[l + 1];
      const uint32_t nrep = c.clustering ? c.rep_off[hi + 1] - c.rep_off[hi] : 0;
      for (uint32_t r = 0; r < nrep; ++r) {
        const uint32_t ri = c.rep_off[hi] + r;
        const uint32_t dpos = c.rep_pos[ri];
        for (uint32_t e = eb; e < ee; ++e) {
          if (ql.entry_qid[e] < qb || ql.entry_qid[e] >= qe) continue;
          st->units_rep += 1;
          const uint32_t qpos = ql.entry_pos[e], qid = ql.entry_qid[e];
          const bool check = ql.entry_check[e] != 0;
          const int d = check ? gzo_distance(qs, qpos, ds, dpos, L, db_.reduced_map) : 0;
          SeedHit sh = {qid, qpos, dpos, l, r, 0, e - eb};
          if (d <= threshold && Passes(qs, qpos, ds, dpos, qid)) hits->push_back(sh);
          if (!check || std::max(d - db_.max_mismatches, 0) <= threshold) {
            for (uint32_t m = c.mem_off[ri]; m < c.mem_off[ri + 1]; ++m) {
              SeedHit mh = sh;
              mh.dpos = c.mem_pos[m];
              mh.member = 1 + (m - c.mem_off[ri]);
              if (Passes(qs, qpos, ds, mh.dpos, qid)) hits->push_back(mh);
            }
          }
        }
      }
      for (uint32_t v = c.val_off[hi]; v < c.val_off[hi + 1]; ++v) {
        const uint32_t dpos = c.val_pos[v];
        for (uint32_t e = eb; e < ee; ++e) {
          if (ql.entry_qid[e] < qb || ql.entry_qid[e] >= qe) continue;
          st->units_value += 1;
          const uint32_t qpos = ql.entry_pos[e], qid = ql.entry_qid[e];
          SeedHit sh = {qid, qpos, dpos, l, nrep + (v - c.val_off[hi]), 0, e - eb};
          if (Passes(qs, qpos, ds, dpos, qid)) hits->push_back(sh);
        }
      }
    }
    // the device returns seeds grouped by query, in no particular order
    std::reverse(hits->begin(), hits->end());
    std::vector<uint32_t> &qstart = qstart_[slot];
    std::vector<SeedHit> &grouped = grouped_[slot];
    qstart.assign(nq + 1, 0);
    for (const SeedHit &h : *hits) ++qstart[h.qid - qb + 1];
    for (uint32_t k = 0; k < nq; ++k) qstart[k + 1] += qstart[k];
    grouped.resize(hits->size());
    std::vector<uint32_t> fill(qstart.begin(), qstart.end() - 1);
    for (const SeedHit &h : *hits) grouped[fill[h.qid - qb]++] = h;
    view->count = grouped.size();
    view->qstart = qstart.data();
    if (slot == 1) {
      view->hits = grouped.data();
      return 0;
    }
    // every other call serves a compact view like the product's: positions,
    // and the records on request, by ranges
    std::vector<HitPosition> &pos = positions_;
    pos.resize(grouped.size());
    for (size_t k = 0; k < grouped.size(); ++k) pos[k] = {grouped[k].qpos, grouped[k].dpos};
    view->hits = nullptr;
    view->positions = pos.data();
    const SeedHit *rec = grouped.data();
    const size_t n_rec = grouped.size();
    view->fetch = [this, rec, n_rec](size_t n, const size_t *begin, const size_t *count,
                                     const SeedHit **records) -> int {
      std::lock_guard<std::mutex> lock(fetch_mutex_);
      fetched_.clear();
      for (size_t i = 0; i < n; ++i) {
        if (begin[i] + count[i] > n_rec) return -3;
        fetched_.insert(fetched_.end(), rec + begin[i], rec + begin[i] + count[i]);
      }
      *records = fetched_.data();
      return 0;
    };
    return 0;
  }

  int GappedExtend(size_t ci, uint32_t nr, uint32_t nf, uint32_t *qpos, uint32_t *dpos,
                   int *scores, uint32_t *cells, const uint16_t *, bool, AlignStats *,
                   const uint32_t *) override {
    const uint8_t *qs = concat_->data();
    const uint8_t *ds = db_.chunks[ci].seq.data();
    for (uint32_t k = 0; k < nr + nf; ++k) {
      const int reverse = k < nr ? 1 : 0;
      int s = 0, p0 = 0, p1 = 0;
      uint32_t ne = 0;
      uint64_t c = 0;
      gzo_gapped_one_side(qs + qpos[k], (uint32_t)concat_->size(), ds + dpos[k], kDelimiter,
                          reverse, p_.matrix, p_.nl, p_.gap_open, p_.gap_extension,
                          presearch_cutoff_, &s, &p0, &p1, nullptr, &ne, &c);
      qpos[k] += p0;
      dpos[k] += p1;
      scores[k] = s;
      if (cells) cells[k] = (uint32_t)c;
    }
    return 0;
  }

  int Tracebacks(size_t ci, const std::vector<TracebackJob> &jobs, int cutoff,
                 std::vector<TracebackOut> *out, AlignStats *) override {
    out->resize(jobs.size());
    for (size_t k = 0; k < jobs.size(); ++k) {
      const TracebackJob &j = jobs[k];
      (*out)[k] = TracebackOneHost(concat_->data() + j.query_offset, j.query_length,
                                   db_.chunks[ci].seq.data(), j.hit_q, j.hit_d, p_, cutoff);
      // exercise the host fall-back paths too
      if (k % 97 == 5) (*out)[k].flags |= kTracebackOverflow;
      if (g_forced_stale > 0 && k % g_forced_stale == 3) (*out)[k].flags |= kTracebackStale;
    }
    return 0;
  }

  void EndQueryChunk() override {}

 private:
  bool Passes(const uint8_t *qs, uint32_t qpos, const uint8_t *ds, uint32_t dpos,
              uint32_t qid) const {
    return gzo_ungapped_with_trigger(qs, qpos, ds, dpos, kDelimiter, p_.matrix, p_.nl,
                                     cut_[qid], trig_[qid]) > trig_[qid];
  }

  const Database &db_;
  const SearchParameters &p_;
  const std::vector<uint8_t> *concat_ = nullptr;
  QueryLists lists_;
  std::vector<int> cut_, trig_;
  int presearch_cutoff_ = 0;
  std::vector<SeedHit> raw_, grouped_[2];
  std::vector<HitPosition> positions_;  // the compact view's (slot 0 only)
  std::vector<SeedHit> fetched_;        // the records of its last fetch
  std::mutex fetch_mutex_;
  std::vector<uint32_t> qstart_[2];
  int flip_ = 0;
};

std::vector<FastaRecord> Records(const char *const *names, const char *residues,
                                 const uint64_t *offsets, uint64_t n) {
  std::vector<FastaRecord> r(n);
  for (uint64_t i = 0; i < n; ++i) {
    r[i].name = names[i];
    r[i].sequence.assign(residues + offsets[i], residues + offsets[i + 1]);
  }
  return r;
}

}  // namespace

extern "C" {

/* The product's AddResults (results.cpp) on the same arguments as
 * gzo_add_results (gz_oracle.h); the subject of an added result comes from
 * DbChunk::SubjectOf over `subject_offsets`. */
int gzh_add_results(const uint32_t *subject_offsets, uint32_t number_subjects, int chunk_id,
                    uint32_t max_results, uint32_t max_results_per_subject,
                    gzo_presearched *added, uint32_t number_added, gzo_presearched *results,
                    uint32_t *number_results) {
  DbChunk c;
  c.offsets.assign(subject_offsets, subject_offsets + number_subjects + 1);
  c.names.assign(number_subjects, std::string());
  SearchParameters p;
  p.max_results = max_results;
  p.max_results_per_subject = max_results_per_subject;
  auto to = [](const gzo_presearched &x) {
    Presearched r;
    r.chunk = x.chunk;
    r.subject = x.subject;
    r.score = x.score;
    r.hit = {x.hit_q, x.hit_d};
    r.start = {x.start_q, x.start_d};
    r.end = {x.end_q, x.end_d};
    r.hit_count = x.hit_count;
    return r;
  };
  std::vector<Presearched> a, r;
  for (uint32_t i = 0; i < number_added; ++i) a.push_back(to(added[i]));
  for (uint32_t i = 0; i < *number_results; ++i) r.push_back(to(results[i]));
  AddResultsScratch scratch;
  AddResults(c, chunk_id, p, scratch, a, r);
  for (size_t i = 0; i < r.size(); ++i) {
    gzo_presearched &x = results[i];
    x.chunk = r[i].chunk;
    x.subject = r[i].subject;
    x.score = r[i].score;
    x.hit_q = r[i].hit.q;
    x.hit_d = r[i].hit.d;
    x.start_q = r[i].start.q;
    x.start_d = r[i].start.d;
    x.end_q = r[i].end.q;
    x.end_d = r[i].end.d;
    x.hit_count = r[i].hit_count;
  }
  *number_results = (uint32_t)r.size();
  return 0;
}

void gzh_force_stale(int every) { g_forced_stale = every; }

/* The product's output number formatter (FormatG6) on n doubles, each text
 * compared with printf's "%g" and with std::to_chars(general, 6).  Returns
 * the number of values where any of the three differ; *first_bad = the
 * first such index (n when none).  out (n * 32 bytes, may be NULL) receives
 * the product's texts, NUL-terminated. */
uint64_t gzh_format_g6_check(const double *v, uint64_t n, char *out, uint64_t *first_bad) {
  uint64_t bad = 0;
  *first_bad = n;
  for (uint64_t i = 0; i < n; ++i) {
    char a[32], b[32], c[32];
    *gz::FormatG6(v[i], a, a + sizeof(a)) = 0;
    snprintf(b, sizeof(b), "%g", v[i]);
    *std::to_chars(c, c + sizeof(c), v[i], std::chars_format::general, 6).ptr = 0;
    if (strcmp(a, b) != 0 || strcmp(a, c) != 0) {
      if (bad++ == 0) *first_bad = i;
    }
    if (out) memcpy(out + i * 32, a, 32);
  }
  return bad;
}

/* The product's host flow with CPU-emulated device stages.  stats[0..5] =
 * seed hits, chained hits, gapped extensions, gapped cells, results,
 * sequential build chunks.  Returns 0 on success; *output is malloc'd. */
int gzh_align(const char *const *qnames, const char *qres, const uint64_t *qoffs, uint64_t nq,
              int translated, int filter, const char *const *dnames, const char *dres,
              const uint64_t *doffs, uint64_t nd, int clustering, uint64_t db_chunk_size,
              const gz_aligner_parameters *prm, char **output, size_t *output_length,
              uint64_t *stats) {
  std::vector<int> m;
  if (!ParseScoreMatrix(Blosum62Text(), m)) return 1;
  QueryOptions qo;
  qo.translated = translated != 0;
  qo.filter = filter != 0;
  QuerySet q;
  BuildQueries(Records(qnames, qres, qoffs, nq), qo, m.data(), kNumberLetters, &q);
  DatabaseOptions dopt;
  dopt.clustering = clustering != 0;
  if (db_chunk_size) dopt.chunk_size = db_chunk_size;
  Database db;
  BuildDatabase(Records(dnames, dres, doffs, nd), dopt, m.data(), kNumberLetters, &db);
  AlignerOptions opt;
  opt.search.matrix = m.data();
  opt.search.nl = kNumberLetters;
  opt.search.gap_open = prm->gap_open;
  opt.search.gap_extension = prm->gap_extension;
  opt.search.normalized_ungapped_cutoff = prm->normalized_ungapped_cutoff;
  opt.search.normalized_gapped_trigger = prm->normalized_gapped_trigger;
  opt.search.normalized_presearch_cutoff = prm->normalized_presearch_cutoff;
  opt.search.normalized_result_cutoff = prm->normalized_result_cutoff;
  opt.search.max_results = prm->max_results;
  opt.search.max_results_per_subject = prm->max_results_per_subject;
  opt.queries_chunk_size = prm->queries_chunk_size ? prm->queries_chunk_size : (1ull << 27);
  opt.threads = prm->threads;
  opt.schedule = prm->schedule == GZ_SCHEDULE_LENGTH_SORTED ? Schedule::kLengthSorted
                                                            : Schedule::kPlain;
  opt.record_cells = prm->record_traces != 0;
  opt.batched = prm->batched != 0;
  opt.pipeline_blocks = prm->pipeline_blocks ? prm->pipeline_blocks : 8;
  CpuDevice dev(db, opt.search);
  std::string out;
  AlignStats st;
  std::vector<GappedTrace> traces;
  const int r = AlignQueries(q, db, opt, &dev, &out, &st, &traces, nullptr);
  if (r != 0) return r;
  char *buf = (char *)malloc(out.size() + 1);
  memcpy(buf, out.data(), out.size());
  buf[out.size()] = 0;
  *output = buf;
  *output_length = out.size();
  if (stats) {
    stats[0] = st.seed_hits;
    stats[1] = st.chained_hits;
    stats[2] = st.gapped_extensions;
    stats[3] = st.gapped_cells;
    stats[4] = st.results;
    stats[5] = st.sequential_build_chunks;
  }
  return 0;
}

void gzh_free(void *p) { free(p); }

/* The product's word-at-a-time seed tests (ungapped.h) run on the host over
 * padded packed arrays, for checks against the oracle; table_bytes selects
 * the byte (1) or 16-bit (2) score tables. */
void gzh_ungapped_words(const uint64_t *qw, const uint64_t *dw, uint64_t n,
                        const uint32_t *qpos, const uint32_t *dpos, const int *matrix,
                        uint32_t nl, const int *cutoff, const int *trigger, int table_bytes,
                        int *out) {
  PairTables<int8_t> t8;
  PairTables<int16_t> t16;
  for (uint32_t i = 0; i < kTableDim * kTableDim; ++i) {
    PairTablesFill(t8, matrix, nl, i);
    PairTablesFill(t16, matrix, nl, i);
  }
  for (uint64_t k = 0; k < n; ++k)
    out[k] = table_bytes == 1
                 ? UngappedWithTrigger(qw, qpos[k], dw, dpos[k], &t8, cutoff[k], trigger[k])
                 : UngappedWithTrigger(qw, qpos[k], dw, dpos[k], &t16, cutoff[k], trigger[k]);
}

void gzh_distance_words(const uint64_t *qw, const uint64_t *dw, uint64_t n,
                        const uint32_t *qc, const uint32_t *dc, uint32_t length,
                        const uint8_t *reduced_map, int *out) {
  uint8_t differs[kTableDim * kTableDim];
  for (uint32_t i = 0; i < kTableDim * kTableDim; ++i) differs[i] = DiffersEntry(reduced_map, i);
  for (uint64_t k = 0; k < n; ++k) {
    out[k] = DistanceWindow(qw, qc[k], dw, dc[k], length, differs);
    // the preloaded-window form the seed kernel uses for length <= 12
    if (length <= kCodesPerWord &&
        DistanceFrom(LoadSeedWindow(qw, qc[k]), LoadSeedWindow(dw, dc[k]), length, differs) !=
            out[k])
      out[k] = -1000;
  }
}

/* The chunk replay in the reference's order (ReplayChunkHost) against the
 * rebuild of only the results that read bytes left by earlier extensions
 * (ResolveStaleChunk) on one database chunk.  Queries: codes with their
 * delimiters and nq + 1 offsets (query i's leading delimiter; the last
 * offset at the closing delimiter); dcodes: the chunk's codes; hits: (query,
 * hit q, hit d) triples, grouped by query.  out[0] = results whose build on
 * a fresh extender read such bytes, out[1] = 1 when ResolveStaleChunk
 * decided them (0: it asked for the replay), out[2] = results of the two
 * builds that differ, out[3] = flagged results whose fresh build differs
 * from the replay's (where the earlier extensions' bytes mattered);
 * results_out (when set): 9 ints per hit as gzo_build_chunk writes them,
 * the product's results after the rebuild. */
int gzh_stale_resolution(const uint8_t *qcodes, const uint64_t *qoff, uint32_t nq,
                         const uint8_t *dcodes, uint32_t dn, const uint32_t *hits, uint32_t nh,
                         const int *matrix, uint32_t nl, int gap_open, int gap_ext, int cutoff,
                         uint64_t *out, int32_t *results_out) {
  gz::QuerySet q;
  q.codes.assign(qcodes, qcodes + qoff[nq] + 1);  // the last query's closing delimiter
  q.offsets.assign(qoff, qoff + nq + 1);
  q.names.assign(nq, "q");
  gz::DbChunk c;
  c.seq.assign(dcodes, dcodes + dn);
  gz::SearchParameters p;
  p.matrix = matrix;
  p.nl = nl;
  p.gap_open = gap_open;
  p.gap_extension = gap_ext;
  std::vector<std::vector<gz::Presearched> > pre(nq);
  for (uint32_t i = 0; i < nh; ++i) {
    gz::Presearched r = {};
    r.chunk = 0;
    r.hit.q = hits[3 * i + 1];
    r.hit.d = hits[3 * i + 2];
    pre[hits[3 * i]].push_back(r);
  }
  const auto ids = gz::ChunkResultIds(pre, 0);
  std::vector<std::vector<gz::AlignmentResult> > want(nq), got(nq);
  for (uint32_t i = 0; i < nq; ++i) {
    want[i].resize(pre[i].size());
    got[i].resize(pre[i].size());
  }
  gz::ReplayChunkHost(q, c, p, cutoff, pre, want, ids);
  std::vector<size_t> flagged;
  for (size_t k = 0; k < ids.size(); ++k)
    if (!gz::BuildOneFresh(q, ids[k].first, c, pre[ids[k].first][ids[k].second], p, cutoff,
                           &got[ids[k].first][ids[k].second]))
      flagged.push_back(k);
  auto same = [](const gz::AlignmentResult &a, const gz::AlignmentResult &b) {
    return a.score == b.score && a.length == b.length && a.mismatches == b.mismatches &&
           a.gap_openings == b.gap_openings && a.gap_extensions == b.gap_extensions &&
           a.start.q == b.start.q && a.start.d == b.start.d && a.end.q == b.end.q &&
           a.end.d == b.end.d;
  };
  uint64_t history = 0;  // flagged results whose fresh build is not the replay's
  for (size_t k : flagged)
    if (!same(got[ids[k].first][ids[k].second], want[ids[k].first][ids[k].second])) ++history;
  const bool ok = gz::ResolveStaleChunk(q, c, p, cutoff, pre, got, ids, flagged);
  uint64_t differ = 0;
  for (uint32_t i = 0; i < nq; ++i)
    for (size_t r = 0; r < want[i].size(); ++r)
      if (!same(want[i][r], got[i][r])) ++differ;
  out[0] = flagged.size();
  out[1] = ok ? 1 : 0;
  out[2] = differ;
  out[3] = history;
  if (results_out) {
    for (size_t k = 0; k < ids.size(); ++k) {
      const gz::AlignmentResult &r = got[ids[k].first][ids[k].second];
      int32_t *o = results_out + 9 * k;
      o[0] = r.score;
      o[1] = (int32_t)r.length;
      o[2] = (int32_t)r.mismatches;
      o[3] = (int32_t)r.gap_openings;
      o[4] = (int32_t)r.gap_extensions;
      o[5] = (int32_t)r.start.q;
      o[6] = (int32_t)r.start.d;
      o[7] = (int32_t)r.end.q;
      o[8] = (int32_t)r.end.d;
    }
  }
  return 0;
}

}  // extern "C"
