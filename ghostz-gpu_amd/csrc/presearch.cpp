// presearch.cpp -- see presearch.h.
#include "presearch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <numeric>
#include <sstream>
#include <thread>

#include "gz_common.h"

namespace gz {

namespace {

double Now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Optional per-block stage trace of the batched pipeline: with the
// environment variable GZ_PIPELINE_TRACE naming a file, one line per stage
// run ("stage block start_ms end_ms", from the trace's creation) is appended
// to it when the trace ends.
class StageTrace {
 public:
  StageTrace() {
    const char *p = getenv("GZ_PIPELINE_TRACE");
    if (p && *p) path_ = p;
    t0_ = Now();
  }
  ~StageTrace() {
    if (path_.empty()) return;
    FILE *f = fopen(path_.c_str(), "a");
    if (!f) return;
    fprintf(f, "# chunk\n");
    for (const Event &e : events_)
      fprintf(f, "%s %zu %.3f %.3f\n", e.stage, e.block, (e.t0 - t0_) * 1e3, (e.t1 - t0_) * 1e3);
    fclose(f);
  }
  void Add(const char *stage, size_t block, double t0, double t1) {
    if (path_.empty()) return;
    std::lock_guard<std::mutex> lock(mu_);
    events_.push_back({stage, block, t0, t1});
  }

 private:
  struct Event {
    const char *stage;
    size_t block;
    double t0, t1;
  };
  std::string path_;
  double t0_ = 0;
  std::mutex mu_;
  std::vector<Event> events_;
};

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

// A passing seed with its one-thread CPU enumeration key (list rank,
// representative or value index, member index + 1 or 0, entry index).
struct KeyedHit {
  uint64_t hi, lo;  // (list << 32 | rv), (member << 32 | entry)
  uint32_t q, d;
  bool operator<(const KeyedHit &o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
};

// (dpos << 32 | index) of n hits sorted by dpos into *keys (LSD radix over
// the bytes of dpos that differ); false when two hits share a dpos.
template <class Hits>
bool SortByPosition(const Hits *hits, uint32_t n, std::vector<uint64_t> *keys,
                    std::vector<uint64_t> *tmp) {
  keys->resize(n);
  tmp->resize(n);
  uint64_t *a = keys->data(), *b = tmp->data();
  uint32_t count[4][256] = {};
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t d = hits[k].dpos;
    a[k] = ((uint64_t)d << 32) | k;
    ++count[0][d & 0xff];
    ++count[1][(d >> 8) & 0xff];
    ++count[2][(d >> 16) & 0xff];
    ++count[3][d >> 24];
  }
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 32 + 8 * pass;
    if (count[pass][(a[0] >> shift) & 0xff] == n) continue;  // one digit value
    uint32_t sum = 0;
    for (int v = 0; v < 256; ++v) {
      const uint32_t c = count[pass][v];
      count[pass][v] = sum;
      sum += c;
    }
    for (uint32_t k = 0; k < n; ++k) b[count[pass][(a[k] >> shift) & 0xff]++] = a[k];
    std::swap(a, b);
  }
  if (a != keys->data()) std::copy(a, a + n, keys->data());
  for (uint32_t k = 1; k < n; ++k)
    if (((*keys)[k] >> 32) == ((*keys)[k - 1] >> 32)) return false;
  return true;
}

#define GZ_STEP(x)            \
  do {                        \
    const int _r = (x);       \
    if (_r != 0) return _r;   \
  } while (0)

int AlignQueryChunk(const QuerySet &q, const Database &db, const AlignerOptions &opt,
                    PresearchDevice *dev, std::string *output, AlignStats *st,
                    std::vector<GappedTrace> *traces, std::vector<uint8_t> &concat) {
  const SearchParameters &P = opt.search;
  const size_t nq = q.size();
  const int threads = opt.threads > 0 ? opt.threads : (int)std::thread::hardware_concurrency();
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
  const double lists_before = st->seconds_query_lists;
  std::vector<uint64_t> coff;
  ConcatenatedOffsets(q, &coff, &concat);
  GZ_STEP(dev->BeginQueryChunk(q, concat, coff, cut, trig, presearch_cutoff, st));
  st->seconds_upload += Now() - t - (st->seconds_query_lists - lists_before);

  std::vector<std::vector<Presearched> > pre(nq);
  std::vector<std::vector<Hit> > chained(nq);
  std::vector<std::vector<AlignmentResult> > results(nq);

  // 2. per-query enumeration order and chain filter of queries [qb, qe)
  //    (hits: the seed view of a run that started at query hb); a failed
  //    copy of a compact view's records is left in chain_error
  std::atomic<int> chain_error(0);
  // the keys (list, rv, member, entry) of a query's hits are distinct, so
  // sorting them gives exactly the one-thread enumeration order
  auto chain_records = [&](const DbChunk &chunk, size_t qi, const SeedHit *rec, uint32_t n,
                           std::vector<KeyedHit> &keyed, std::vector<uint64_t> &scratch) {
    keyed.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
      const SeedHit &h = rec[k];
      keyed[k].hi = ((uint64_t)h.list << 32) | h.rv;
      keyed[k].lo = ((uint64_t)h.member << 32) | h.entry;
      keyed[k].q = h.qpos - (uint32_t)coff[qi];
      keyed[k].d = h.dpos;
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<Hit> &hv = chained[qi];
    hv.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
      hv[k].q = keyed[k].q;
      hv[k].d = keyed[k].d;
    }
    ChainFilter(q.Sequence(qi), q.Length(qi), cut[qi], chunk.seq.data(), P.matrix, P.nl, hv,
                scratch);
  };
  auto chain = [&](const DbChunk &chunk, const SeedHitView &hits, size_t hb, size_t qb,
                   size_t qe) {
    // queries with two hits at one database position whose records are
    // still on the device (a compact view)
    std::vector<uint8_t> tied(qe - qb, 0);
    ParallelBlocks(qe - qb, threads, 64, [&](size_t b, size_t e) {
      std::vector<uint64_t> scratch, by_d, by_d_tmp;
      std::vector<KeyedHit> keyed;
      for (size_t qi = qb + b; qi < qb + e; ++qi) {
        const uint32_t s0 = hits.qstart[qi - hb], s1 = hits.qstart[qi - hb + 1];
        if (s0 == s1) continue;
        // ChainFilter sorts by database position with an unstable sort.
        // When no two hits share a position that order is unique, and the
        // enumeration order it starts from does not matter.
        const bool unique = hits.hits ? SortByPosition(hits.hits + s0, s1 - s0, &by_d, &by_d_tmp)
                                      : SortByPosition(hits.positions + s0, s1 - s0, &by_d,
                                                       &by_d_tmp);
        if (unique) {
          std::vector<Hit> &hv = chained[qi];
          hv.resize(s1 - s0);
          for (uint32_t k = 0; k < s1 - s0; ++k) {
            const uint32_t x = s0 + (uint32_t)by_d[k];
            hv[k].q = (hits.hits ? hits.hits[x].qpos : hits.positions[x].qpos) -
                      (uint32_t)coff[qi];
            hv[k].d = hits.hits ? hits.hits[x].dpos : hits.positions[x].dpos;
          }
          ChainFilterSorted(q.Sequence(qi), q.Length(qi), cut[qi], chunk.seq.data(), P.matrix,
                            P.nl, hv, scratch);
        } else if (hits.hits) {
          chain_records(chunk, qi, hits.hits + s0, s1 - s0, keyed, scratch);
        } else {
          tied[qi - qb] = 1;
        }
      }
    });
    if (hits.hits) return;
    // the tied queries' records come from the device in one fetch (adjacent
    // queries' ranges merged), then are ordered like above
    std::vector<size_t> begin, count, at, which;
    size_t total = 0;
    for (size_t qi = qb; qi < qe; ++qi) {
      if (!tied[qi - qb]) continue;
      const size_t s0 = hits.qstart[qi - hb], s1 = hits.qstart[qi - hb + 1];
      which.push_back(qi);
      at.push_back(total);
      if (!begin.empty() && begin.back() + count.back() == s0) count.back() += s1 - s0;
      else {
        begin.push_back(s0);
        count.push_back(s1 - s0);
      }
      total += s1 - s0;
    }
    if (total == 0) return;
    const SeedHit *fetched = nullptr;
    const int f = hits.fetch(begin.size(), begin.data(), count.data(), &fetched);
    if (f != 0) {
      chain_error.store(f);
      return;
    }
    ParallelBlocks(which.size(), threads, 4, [&](size_t b, size_t e) {
      std::vector<uint64_t> scratch;
      std::vector<KeyedHit> keyed;
      for (size_t k = b; k < e; ++k) {
        const size_t qi = which[k];
        chain_records(chunk, qi, fetched + at[k], hits.qstart[qi - hb + 1] - hits.qstart[qi - hb],
                      keyed, scratch);
      }
    });
  };
  // 3. gapped extension of both sides of the chained hits of [qb, qe)
  struct Block {
    size_t qb = 0, qe = 0, slot = 0;
    std::vector<uint32_t> gstart;  // per query: its first job; [n] = all jobs
    uint32_t bucket_off[kLengthBuckets + 1] = {0, 0, 0, 0, 0};
    // jobs [0, G) reverse sides, [G, 2G) forward sides: positions in,
    // extension ends and scores out
    uint32_t *aq = nullptr, *ad = nullptr;
    int *sc = nullptr;
    uint32_t G = 0;
    std::vector<uint16_t> keys;  // query side lengths (ordered schedules)
  };
  // the order the device runs the jobs in never changes their results
  const bool learned = opt.schedule == Schedule::kLearned && opt.cost_model != nullptr;
  // 3a. the jobs of a block laid out and handed to the device (pipeline
  // slot `slot`), which may start ordering them
  auto gapped_prepare = [&](size_t ci, Block &B, int slot) -> int {
    const size_t n = B.qe - B.qb;
    B.gstart.assign(n + 1, 0);
    if (opt.batched) {
      // the batched form lays each side's jobs out by query-length bucket
      // (queries in order inside a bucket), so that the device can run every
      // bucket with its own tuned configuration
      uint32_t count[kLengthBuckets] = {0, 0, 0, 0};
      for (size_t k = 0; k < n; ++k)
        count[LengthBucket(q.Length(B.qb + k))] += (uint32_t)chained[B.qb + k].size();
      for (int b = 0; b < kLengthBuckets; ++b) B.bucket_off[b + 1] = B.bucket_off[b] + count[b];
      uint32_t cursor[kLengthBuckets];
      for (int b = 0; b < kLengthBuckets; ++b) cursor[b] = B.bucket_off[b];
      for (size_t k = 0; k < n; ++k) {
        const int b = LengthBucket(q.Length(B.qb + k));
        B.gstart[k] = cursor[b];
        cursor[b] += (uint32_t)chained[B.qb + k].size();
      }
      B.gstart[n] = B.bucket_off[kLengthBuckets];
    } else {
      for (size_t k = 0; k < n; ++k)
        B.gstart[k + 1] = B.gstart[k] + (uint32_t)chained[B.qb + k].size();
      for (int b = 1; b <= kLengthBuckets; ++b) B.bucket_off[b] = B.gstart[n];
    }
    const uint32_t G = B.gstart[n];
    B.G = G;
    if (G == 0) return 0;
    const size_t jobs = 2 * (size_t)G;
    uint8_t *buf = dev->JobBuffer(B.slot, jobs * 12);
    B.aq = reinterpret_cast<uint32_t *>(buf);
    B.ad = B.aq + jobs;
    B.sc = reinterpret_cast<int *>(B.ad + jobs);
    // query side lengths: the length-sorted schedule's keys, and the length
    // bins of the learned schedule
    std::vector<uint16_t> &keys = B.keys;
    keys.clear();
    if (opt.schedule == Schedule::kLengthSorted || opt.schedule == Schedule::kLearned)
      keys.resize(jobs);
    ParallelBlocks(n, opt.batched ? std::max(1, threads / 4) : threads, 256,
                   [&](size_t b, size_t e) {
      for (size_t k = b; k < e; ++k) {
        const size_t qi = B.qb + k;
        for (uint32_t j = 0; j < chained[qi].size(); ++j) {
          const uint32_t g = B.gstart[k] + j;
          const Hit &h = chained[qi][j];
          B.aq[g] = (uint32_t)coff[qi] + h.q - 1;  // reverse side starts before the hit
          B.ad[g] = h.d - 1;
          B.aq[G + g] = (uint32_t)coff[qi] + h.q;
          B.ad[G + g] = h.d;
          if (!keys.empty()) {
            // ExtensionSeedList order: by the query distance the side can extend
            const uint32_t r = q.DistanceFromStart(qi, h.q - 1);
            const uint32_t f = q.DistanceFromEnd(qi, h.q);
            keys[g] = (uint16_t)std::min<uint32_t>(r, kOrderKeyLimit - 1);
            keys[G + g] = (uint16_t)std::min<uint32_t>(f, kOrderKeyLimit - 1);
          }
        }
      }
    });
    return dev->GappedPrepare(slot, ci, G, G, B.aq, B.ad, keys.empty() ? nullptr : keys.data(),
                              learned, opt.batched ? B.bucket_off : nullptr);
  };
  // 3b. the block's gapped extensions: extension ends and scores back in
  // the block's arrays
  std::mutex gapped_mu;
  auto gapped_run = [&](size_t ci, Block &B, int slot) -> int {
    const size_t n = B.qe - B.qb;
    const uint32_t G = B.G;
    if (G == 0) return 0;
    const size_t jobs = 2 * (size_t)G;
    std::vector<uint32_t> cells(opt.record_cells ? jobs : 0);
    const int r = dev->GappedRun(slot, ci, G, G, B.aq, B.ad, B.sc,
                                 cells.empty() ? nullptr : cells.data(), st,
                                 opt.batched ? B.bucket_off : nullptr);
    if (r != 0) return r;
    // the two pipeline slots' extensions may finish on two threads
    std::lock_guard<std::mutex> lock(gapped_mu);
    if (!cells.empty()) {
      for (int side = 0; side < 2; ++side)
        for (size_t k = 0; k < n; ++k) {
          const size_t qi = B.qb + k;
          for (uint32_t j = 0; j < chained[qi].size(); ++j) {
            const Hit &h = chained[qi][j];
            GappedTrace tr;
            tr.chunk = (uint32_t)ci;
            tr.reverse = side == 0 ? 1 : 0;
            tr.qpos = (uint32_t)coff[qi] + h.q - (side == 0 ? 1 : 0);
            tr.dpos = h.d - (side == 0 ? 1 : 0);
            tr.cells = cells[(size_t)side * G + B.gstart[k] + j];
            if (traces) traces->push_back(tr);
            st->gapped_cells += tr.cells;
          }
        }
    }
    st->gapped_extensions += jobs;
    st->chained_hits += G;
    return 0;
  };
  // 4. presearched results and AddResults for the queries of a block
  auto add = [&](size_t ci, const Block &B) {
    const DbChunk &chunk = db.chunks[ci];
    ParallelBlocks(B.qe - B.qb, threads, 64, [&](size_t b, size_t e) {
      AddResultsScratch scratch;
      std::vector<Presearched> tmp;
      for (size_t k = b; k < e; ++k) {
        const size_t qi = B.qb + k;
        tmp.clear();
        for (uint32_t j = 0; j < chained[qi].size(); ++j) {
          const uint32_t g = B.gstart[k] + j;
          const Hit &h = chained[qi][j];
          Presearched r;
          r.chunk = -1;
          r.subject = 0;
          r.hit.q = h.q;
          r.hit.d = h.d;
          r.start.q = B.aq[g] - (uint32_t)coff[qi];
          r.start.d = B.ad[g];
          r.end.q = B.aq[B.G + g] - (uint32_t)coff[qi];
          r.end.d = B.ad[B.G + g];
          r.score = B.sc[g] + B.sc[B.G + g];
          r.hit_count = 1;
          tmp.push_back(r);
        }
        if (!tmp.empty() || !pre[qi].empty())
          AddResults(chunk, (int)ci, P, scratch, tmp, pre[qi]);
        std::vector<Hit>().swap(chained[qi]);
      }
    });
  };
  // 5. result build of queries [qb, qe), device part: traceback jobs of every
  //    database chunk (kTracebackUnsupported when the device has none)
  struct BuildPart {
    size_t ci;
    std::vector<std::pair<uint32_t, uint32_t> > ids;
    std::vector<TracebackOut> outs;
  };
  auto build_device = [&](size_t qb, size_t qe, std::vector<BuildPart> *parts) -> int {
    parts->clear();
    for (size_t i = qb; i < qe; ++i) results[i].assign(pre[i].size(), AlignmentResult());
    for (size_t ci = 0; ci < db.chunks.size(); ++ci) {
      BuildPart part;
      part.ci = ci;
      part.ids = ChunkResultIds(pre, ci, qb, qe);
      if (part.ids.empty()) continue;
      std::vector<TracebackJob> jobs(part.ids.size());
      for (size_t k = 0; k < part.ids.size(); ++k) {
        const uint32_t qi = part.ids[k].first;
        const Presearched &pr = pre[qi][part.ids[k].second];
        jobs[k].query_offset = (uint32_t)coff[qi];
        jobs[k].query_length = q.Length(qi);
        jobs[k].hit_q = pr.hit.q;
        jobs[k].hit_d = pr.hit.d;
      }
      const int rc = dev->Tracebacks(ci, jobs, result_cutoff, &part.outs, st);
      if (rc != 0) return rc;
      parts->push_back(std::move(part));
    }
    return 0;
  };
  //    host part: copy the device's results and rebuild on the host those
  //    the device left (a traceback that outgrew its arena); results whose
  //    traceback read a byte an earlier extension left are rebuilt with the
  //    chunk's earlier extensions in the reference order when `whole` (the
  //    parts hold every query of the chunk), else their chunk is listed in
  //    stale_chunks (to be rebuilt whole); so is a chunk that rebuild
  //    cannot decide.  host_jobs counts results redone on the host
  auto build_host = [&](std::vector<BuildPart> &parts, uint64_t *host_jobs,
                        std::vector<size_t> *stale_chunks, bool whole) {
    for (BuildPart &part : parts) {
      const DbChunk &chunk = db.chunks[part.ci];
      const auto &ids = part.ids;
      const auto &outs = part.outs;
      std::vector<size_t> redo, flagged;
      for (size_t k = 0; k < ids.size(); ++k) {
        if (outs[k].flags & kTracebackStale) flagged.push_back(k);
        else if (outs[k].flags & kTracebackOverflow) redo.push_back(k);
      }
      *host_jobs += redo.size();
      ParallelBlocks(ids.size(), threads, 4096, [&](size_t b, size_t e) {
        for (size_t k = b; k < e; ++k) {
          if (outs[k].flags) continue;
          const Presearched &pr = pre[ids[k].first][ids[k].second];
          const TracebackOut &o = outs[k];
          AlignmentResult &r = results[ids[k].first][ids[k].second];
          r.chunk = pr.chunk;
          r.subject = pr.subject;
          r.score = o.score;
          r.length = o.length;
          r.mismatches = o.mismatches;
          r.gap_openings = o.gap_openings;
          r.gap_extensions = o.gap_extensions;
          r.hit = pr.hit;
          r.start.q = o.start_q;
          r.start.d = o.start_d;
          r.end.q = o.end_q;
          r.end.d = o.end_d;
        }
      });
      // results whose traceback outgrew a lane's arena, each rebuilt on its
      // own (independent results: any thread, any order)
      std::vector<uint8_t> redo_stale(redo.size(), 0);
      ParallelBlocks(redo.size(), threads, 1, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
          const size_t k = redo[i];
          const uint32_t qi = ids[k].first, r = ids[k].second;
          if (!BuildOneFresh(q, qi, chunk, pre[qi][r], P, result_cutoff, &results[qi][r]))
            redo_stale[i] = 1;
        }
      });
      for (size_t i = 0; i < redo.size(); ++i)
        if (redo_stale[i]) flagged.push_back(redo[i]);
      if (flagged.empty()) continue;
      std::sort(flagged.begin(), flagged.end());
      if (!whole || !ResolveStaleChunk(q, chunk, P, result_cutoff, pre, results, ids, flagged))
        stale_chunks->push_back(part.ci);
    }
  };
  //    the whole query chunk, with the reference-order replay of every
  //    database chunk where a traceback read an earlier extension's byte
  auto build_all = [&]() -> int {
    std::vector<BuildPart> parts;
    const int rc = build_device(0, nq, &parts);
    if (rc == PresearchDevice::kTracebackUnsupported) {
      for (size_t i = 0; i < nq; ++i) results[i].assign(pre[i].size(), AlignmentResult());
      for (size_t ci = 0; ci < db.chunks.size(); ++ci)
        BuildChunkHost(q, db.chunks[ci], P, result_cutoff, pre, results,
                       ChunkResultIds(pre, ci), threads, &st->sequential_build_chunks);
    } else {
      GZ_STEP(rc);
      std::vector<size_t> stale;
      build_host(parts, &st->traceback_host_jobs, &stale, true);
      for (size_t ci : stale) {
        ++st->sequential_build_chunks;
        ReplayChunkHost(q, db.chunks[ci], P, result_cutoff, pre, results,
                        ChunkResultIds(pre, ci));
      }
    }
    SortResults(results, 0, nq);
    return 0;
  };

  if (!opt.batched) {
    // plain: each stage over all queries, one after the other
    for (size_t ci = 0; ci < db.chunks.size(); ++ci) {
      const DbChunk &chunk = db.chunks[ci];
      t = Now();
      SeedHitView hits;
      GZ_STEP(dev->SeedHits(ci, 0, (uint32_t)nq, &hits, st));
      st->seed_hits += hits.count;
      st->seconds_seed += Now() - t;
      t = Now();
      chain(chunk, hits, 0, 0, nq);
      GZ_STEP(chain_error.load());
      st->seconds_chain += Now() - t;
      t = Now();
      Block all;
      all.qb = 0;
      all.qe = nq;
      GZ_STEP(gapped_prepare(ci, all, 0));
      GZ_STEP(gapped_run(ci, all, 0));
      st->seconds_gapped += Now() - t;
      t = Now();
      add(ci, all);
      st->seconds_add += Now() - t;
    }
    t = Now();
    GZ_STEP(build_all());
    dev->EndQueryChunk();
    for (auto &r : results) st->results += r.size();
    st->seconds_build += Now() - t;
    t = Now();
    WriteOutputText(output, q, db, results, 0, nq, threads);
    st->seconds_output += Now() - t;
    return 0;
  }

  // batched: blocks of queries (multiples of 64, equal residues) flow through
  // seed (device) -> chain filter (host) -> gapped extension (device) ->
  // AddResults (host), and after the last database chunk through the result
  // build (device, then host) and the output text; one host thread drives
  // the device, so the host stages of one block overlap the device stages
  // of the next
  std::vector<Block> blocks;
  {
    const size_t want = std::max<size_t>(1, std::min<size_t>(opt.pipeline_blocks,
                                                             (nq + 63) / 64));
    uint64_t total = 0;
    for (size_t i = 0; i < nq; ++i) total += q.Length(i);
    // block weights ramp up over the first blocks and down over the last
    // ones, so that the pipeline fills and drains on small blocks
    std::vector<double> cum(want + 1, 0.0);
    for (size_t b = 0; b < want; ++b)
      cum[b + 1] = cum[b] + (double)std::min<size_t>(std::min(b + 1, want - b), 4);
    size_t qb = 0;
    uint64_t sum = 0;
    for (size_t b = 0; b < want && qb < nq; ++b) {
      size_t qe = nq;
      if (b + 1 < want) {
        const uint64_t target = (uint64_t)((double)total * cum[b + 1] / cum[want]);
        qe = qb;
        while (qe < nq) {
          sum += q.Length(qe);
          ++qe;
          if (qe % 64 == 0 && sum >= target) break;
        }
      }
      Block B;
      B.qb = qb;
      B.qe = qe;
      B.slot = blocks.size();
      blocks.push_back(B);
      qb = qe;
    }
  }
  const size_t nb = blocks.size();
  StageTrace trace;
  const size_t output_before = output->size();
  std::vector<std::vector<BuildPart> > block_parts(nb);
  std::vector<size_t> stale_chunks;
  bool rebuild = false;  // a block could not be built on its own
  double sec_seed = 0, sec_schedule = 0, sec_gapped = 0, sec_chain = 0, sec_add = 0, sec_build = 0,
         sec_build_host = 0, sec_output = 0;
  uint64_t seed_hits = 0, host_jobs = 0;
  // The pipeline runs over every (database chunk, block) pair as one
  // sequence of tasks t = chunk * nb + block, so that the next chunk's
  // seeding overlaps the previous chunk's last extensions instead of the
  // pipeline draining at every chunk.  A block's host arrays (hits per query,
  // job buffers) are reused by the same block of the next chunk, so its
  // chain filter and gapped extension wait for the previous chunk's
  // AddResults of the block; AddResults runs tasks in order, which keeps
  // every query's merges in chunk order.  The result build and the output
  // follow the last chunk's AddResults.
  const size_t nchunks = db.chunks.size();
  const size_t ntasks = nchunks * nb;
  const size_t last_base = (nchunks - 1) * nb;
  std::vector<SeedHitView> views(ntasks);
  std::mutex mu;
  std::condition_variable cv;
  size_t seeded = 0, chained_upto = 0, prepared = 0, gapped_upto = 0, added = 0, traced = 0;
  int status = 0;
  bool unsupported = false;
  auto wait_for = [&](const size_t &counter, size_t value) {
    std::unique_lock<std::mutex> lock(mu);
    cv.wait(lock, [&] { return counter >= value || status != 0; });
    return status == 0;
  };
  auto publish = [&](size_t &counter, size_t value, int r) {
    {
      std::lock_guard<std::mutex> lock(mu);
      counter = value;
      if (r != 0 && status == 0) status = r;
    }
    cv.notify_all();
  };
  // the seeder and the extender drive the device's two streams
  std::thread seeder([&]() {
    for (size_t t = 0; t < ntasks; ++t) {
      const size_t ci = t / nb, k = t % nb;
      // the view of task t reuses the buffer of task t - 2
      if (t >= 2 && !wait_for(chained_upto, t - 1)) return;
      const double t0 = Now();
      const int r = dev->SeedHits(ci, (uint32_t)blocks[k].qb, (uint32_t)blocks[k].qe,
                                  &views[t], st);
      seed_hits += views[t].count;
      sec_seed += Now() - t0;
      trace.Add("seed", t, t0, Now());
      publish(seeded, t + 1, r);
      if (r != 0) return;
    }
  });
  // the scheduler lays task t's jobs out and hands them to the device,
  // which orders them on its own stream while the extender's task t - 1
  // runs; the extender, the scheduler and the builder drive their own
  // device streams, and their host work overlaps.  Task t uses the device's
  // pipeline slot t % 2, free once task t - 2 has been extended.
  std::thread scheduler([&]() {
    for (size_t t = 0; t < ntasks; ++t) {
      if (!wait_for(chained_upto, t + 1)) return;
      if (t >= 2 && !wait_for(gapped_upto, t - 1)) return;
      const double t0 = Now();
      const int r = gapped_prepare(t / nb, blocks[t % nb], (int)(t % 2));
      sec_schedule += Now() - t0;
      trace.Add("schedule", t, t0, Now());
      publish(prepared, t + 1, r);
      if (r != 0) return;
    }
  });
  // With queries longer than 2048 residues (a long-tailed query set), two
  // extenders run, one per pipeline slot (even and odd tasks), each on its
  // own device stream, so that one task's longest bands overlap the next
  // task's extensions; else one extender runs the tasks in turn (measured:
  // short reads lose about 4% with two, long-tailed queries gain about a
  // third).  Tasks are published as extended in order.
  size_t longest = 0;
  for (size_t i = 0; i < nq; ++i) longest = std::max<size_t>(longest, q.Length(i));
  const size_t stride = longest > 2048 ? 2 : 1;
  dev->SetConcurrentExtensions(stride == 2);
  std::vector<char> extended(ntasks, 0);
  double sec_gapped_slot[2] = {0, 0};
  auto extend_slot = [&](size_t first) {
    for (size_t t = first; t < ntasks; t += stride) {
      if (!wait_for(prepared, t + 1)) return;
      const double t0 = Now();
      const int r = gapped_run(t / nb, blocks[t % nb], (int)(t % 2));
      sec_gapped_slot[first] += Now() - t0;
      trace.Add("gapped", t, t0, Now());
      {
        std::lock_guard<std::mutex> lock(mu);
        extended[t] = 1;
        size_t upto = gapped_upto;
        while (upto < ntasks && extended[upto]) ++upto;
        gapped_upto = upto;
        if (r != 0 && status == 0) status = r;
      }
      cv.notify_all();
      if (r != 0) return;
    }
  };
  std::thread extender([&]() { extend_slot(0); });
  std::thread extender_odd([&]() {
    if (stride == 2) extend_slot(1);
  });
  std::thread builder([&]() {
    for (size_t k = 0; k < nb; ++k) {
      if (!wait_for(added, last_base + k + 1)) return;
      const double t0 = Now();
      int r = 0;
      if (!unsupported) {
        r = build_device(blocks[k].qb, blocks[k].qe, &block_parts[k]);
        if (r == PresearchDevice::kTracebackUnsupported) {
          unsupported = true;
          r = 0;
        }
      }
      sec_build += Now() - t0;
      trace.Add("build_device", k, t0, Now());
      publish(traced, k + 1, r);
      if (r != 0) return;
    }
  });
  std::thread adder([&]() {
    for (size_t t = 0; t < ntasks; ++t) {
      if (!wait_for(gapped_upto, t + 1)) return;
      const double t0 = Now();
      add(t / nb, blocks[t % nb]);
      sec_add += Now() - t0;
      trace.Add("add", t, t0, Now());
      publish(added, t + 1, 0);
    }
  });
  std::thread writer([&]() {
    for (size_t b = 0; b < nb; ++b) {
      if (!wait_for(traced, b + 1)) return;
      if (unsupported || rebuild) continue;
      const double t0 = Now();
      const size_t before = stale_chunks.size();
      build_host(block_parts[b], &host_jobs, &stale_chunks, false);
      std::vector<BuildPart>().swap(block_parts[b]);
      if (stale_chunks.size() != before) {
        rebuild = true;
        continue;
      }
      SortResults(results, blocks[b].qb, blocks[b].qe);
      const double t1 = Now();
      sec_build_host += t1 - t0;
      trace.Add("build_host", b, t0, t1);
      // blocks are written in query order, straight into the output
      WriteOutputText(output, q, db, results, blocks[b].qb, blocks[b].qe, threads);
      // reserve for the whole query chunk, extrapolated from the text so
      // far, so that the text moves at most once more rather than at
      // every doubling
      if (b + 1 < nb) {
        const double share = (double)blocks[b].qe / (double)nq;
        const size_t want =
            output_before + (size_t)((double)(output->size() - output_before) / share * 1.15);
        if (output->capacity() < want) output->reserve(want);
      }
      sec_output += Now() - t1;
      trace.Add("output", b, t1, Now());
    }
  });
  for (size_t t = 0; t < ntasks; ++t) {
    if (!wait_for(seeded, t + 1)) break;
    // the block's hit lists and job arrays are free once the previous
    // chunk's AddResults of the block is done
    if (t >= nb && !wait_for(added, t - nb + 1)) break;
    const size_t ci = t / nb, k = t % nb;
    const double t0 = Now();
    chain(db.chunks[ci], views[t], blocks[k].qb, blocks[k].qb, blocks[k].qe);
    sec_chain += Now() - t0;
    trace.Add("chain", t, t0, Now());
    const int rc = chain_error.load();
    publish(chained_upto, t + 1, rc);
    if (rc != 0) break;
  }
  seeder.join();
  scheduler.join();
  extender.join();
  extender_odd.join();
  sec_gapped = sec_gapped_slot[0] + sec_gapped_slot[1];
  builder.join();
  adder.join();
  writer.join();
  GZ_STEP(status);
  if (unsupported || rebuild) {
    // rebuild every result in the reference's order and write all queries
    t = Now();
    GZ_STEP(build_all());
    sec_build += Now() - t;
    t = Now();
    output->resize(output_before);
    WriteOutputText(output, q, db, results, 0, nq, threads);
    sec_output += Now() - t;
  }
  dev->EndQueryChunk();
  st->seed_hits += seed_hits;
  st->seconds_seed += sec_seed;
  st->seconds_chain += sec_chain;
  st->seconds_gapped += sec_gapped + sec_schedule;
  st->seconds_add += sec_add;
  st->seconds_build += sec_build + sec_build_host;
  st->seconds_output += sec_output;
  st->traceback_host_jobs += host_jobs;
  for (auto &r : results) st->results += r.size();
  return 0;
}

}  // namespace

int AlignQueries(const QuerySet &queries, const Database &db, const AlignerOptions &opt,
                 PresearchDevice *dev, std::string *output, AlignStats *stats,
                 std::vector<GappedTrace> *traces, std::vector<uint8_t> *concat) {
  AlignStats dummy;
  AlignStats *st = stats ? stats : &dummy;
  std::vector<uint8_t> local_concat;
  std::vector<uint8_t> &cc = concat ? *concat : local_concat;
  if (traces) traces->clear();
  const double t0 = Now();
  // query chunks (Queries::SetQueries, src/queries.cpp:183-224)
  const size_t n = queries.size();
  size_t i = 0;
  while (i < n) {
    uint64_t sum = 0;
    size_t j = i;
    while (j < n) {
      const uint64_t next = sum + queries.Length(j);
      if (next >= opt.queries_chunk_size) break;
      sum = next;
      ++j;
    }
    if (j == i) j = i + 1;
    if (i == 0 && j == n) {
      // one chunk holds every query: no copy
      GZ_STEP(AlignQueryChunk(queries, db, opt, dev, output, st, traces, cc));
      ++st->query_chunks;
      break;
    }
    QuerySet part;
    part.translated = queries.translated;
    part.max_length = 0;
    part.offsets.push_back(0);
    for (size_t k = i; k < j; ++k) {
      part.names.push_back(queries.names[k]);
      part.codes.insert(part.codes.end(), queries.Sequence(k),
                        queries.Sequence(k) + queries.Length(k));
      part.offsets.push_back(part.codes.size());
      part.lambda.push_back(queries.lambda[k]);
      part.K.push_back(queries.K[k]);
      if (queries.translated) part.dna_lengths.push_back(queries.dna_lengths[k]);
      part.max_length = std::max(part.max_length, queries.Length(k));
    }
    GZ_STEP(AlignQueryChunk(part, db, opt, dev, output, st, traces, cc));
    ++st->query_chunks;
    i = j;
  }
  st->seconds_total = Now() - t0;
  return 0;
}

}  // namespace gz
