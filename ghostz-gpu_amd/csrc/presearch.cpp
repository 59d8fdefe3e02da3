// presearch.cpp -- see presearch.h.
#include "presearch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
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

// the one-thread CPU enumeration order of passing seeds within one query
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
  std::vector<uint64_t> coff;
  ConcatenatedOffsets(q, &coff, &concat);
  QueryLists lists;
  BuildQueryLists(q, coff, db, threads, &lists);
  st->seconds_query_lists += Now() - t;
  st->lists += lists.NumberLists();

  t = Now();
  GZ_STEP(dev->BeginQueryChunk(q, concat, lists, cut, trig, presearch_cutoff, st));
  st->seconds_upload += Now() - t;

  std::vector<std::vector<Presearched> > pre(nq);
  for (size_t ci = 0; ci < db.chunks.size(); ++ci) {
    const DbChunk &chunk = db.chunks[ci];
    // 1. seed stage
    t = Now();
    SeedHitView hits;
    GZ_STEP(dev->SeedHits(ci, (uint32_t)nq, &hits, st));
    st->seed_hits += hits.count;
    st->seconds_seed += Now() - t;

    std::vector<std::vector<Hit> > chained(nq);
    // 2. per-query enumeration order and chain filter of queries [qb, qe)
    auto chain = [&](size_t qb, size_t qe) {
      ParallelBlocks(qe - qb, threads, 64, [&](size_t b, size_t e) {
        std::vector<uint64_t> scratch;
        std::vector<uint32_t> byq;
        KeyOrder order{hits.hits};
        for (size_t qi = qb + b; qi < qb + e; ++qi) {
          const uint32_t s0 = hits.qstart[qi], s1 = hits.qstart[qi + 1];
          if (s0 == s1) continue;
          byq.resize(s1 - s0);
          std::iota(byq.begin(), byq.end(), s0);
          std::sort(byq.begin(), byq.end(), order);
          std::vector<Hit> &hv = chained[qi];
          hv.resize(s1 - s0);
          for (uint32_t k = 0; k < s1 - s0; ++k) {
            const SeedHit &h = hits.hits[byq[k]];
            hv[k].q = h.qpos - (uint32_t)coff[qi];
            hv[k].d = h.dpos;
          }
          ChainFilter(q.Sequence(qi), q.Length(qi), cut[qi], chunk.seq.data(), P.matrix, P.nl,
                      hv, scratch);
        }
      });
    };
    // 3. gapped extension of both sides of the chained hits of [qb, qe)
    struct Block {
      size_t qb = 0, qe = 0, slot = 0;
      std::vector<uint32_t> gstart;
      // jobs [0, G) reverse sides, [G, 2G) forward sides: positions in,
      // extension ends and scores out
      uint32_t *aq = nullptr, *ad = nullptr;
      int *sc = nullptr;
      uint32_t G = 0;
    };
    auto gapped = [&](Block &B) -> int {
      const size_t n = B.qe - B.qb;
      B.gstart.assign(n + 1, 0);
      for (size_t k = 0; k < n; ++k)
        B.gstart[k + 1] = B.gstart[k] + (uint32_t)chained[B.qb + k].size();
      const uint32_t G = B.gstart[n];
      B.G = G;
      if (G == 0) return 0;
      const size_t jobs = 2 * (size_t)G;
      uint8_t *buf = dev->JobBuffer(B.slot, jobs * 12);
      B.aq = reinterpret_cast<uint32_t *>(buf);
      B.ad = B.aq + jobs;
      B.sc = reinterpret_cast<int *>(B.ad + jobs);
      std::vector<uint16_t> keys;
      if (opt.schedule == Schedule::kLengthSorted) keys.resize(jobs);
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
      // the order the device runs the jobs in never changes their results
      const bool learned = opt.schedule == Schedule::kLearned && opt.cost_model != nullptr;
      std::vector<uint32_t> cells(opt.record_cells ? jobs : 0);
      const int r = dev->GappedExtend(ci, G, G, B.aq, B.ad, B.sc,
                                      cells.empty() ? nullptr : cells.data(),
                                      keys.empty() ? nullptr : keys.data(), learned, st);
      if (r != 0) return r;
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
    auto add = [&](const Block &B) {
      ParallelBlocks(B.qe - B.qb, threads, 64, [&](size_t b, size_t e) {
        std::vector<std::vector<AlignmentPosition> > starts;
        std::vector<Presearched> tmp;
        for (size_t k = b; k < e; ++k) {
          const size_t qi = B.qb + k;
          tmp.clear();
          for (uint32_t g = B.gstart[k]; g < B.gstart[k + 1]; ++g) {
            const Hit &h = chained[qi][g - B.gstart[k]];
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
            AddResults(chunk, (int)ci, P, starts, tmp, pre[qi]);
          std::vector<Hit>().swap(chained[qi]);
        }
      });
    };

    if (!opt.batched) {
      // plain: each stage over all queries, one after the other
      t = Now();
      chain(0, nq);
      st->seconds_chain += Now() - t;
      t = Now();
      Block all;
      all.qb = 0;
      all.qe = nq;
      GZ_STEP(gapped(all));
      st->seconds_gapped += Now() - t;
      t = Now();
      add(all);
      st->seconds_add += Now() - t;
      continue;
    }
    // batched: blocks of queries flow through chain (host) -> gapped (device,
    // on its own host thread) -> AddResults (host), so that the host stages
    // of one block overlap the device stage of the next
    const size_t nblocks = std::max<size_t>(1, std::min<size_t>(opt.pipeline_blocks, nq));
    std::vector<Block> blocks(nblocks);
    {
      // equal shares of seed hits per block
      const uint64_t total = hits.count;
      size_t qb = 0;
      for (size_t b = 0; b < nblocks; ++b) {
        const uint64_t target = total * (b + 1) / nblocks;
        size_t qe = qb;
        while (qe < nq && (b + 1 == nblocks || hits.qstart[qe] < target)) ++qe;
        if (b + 1 == nblocks) qe = nq;
        blocks[b].qb = qb;
        blocks[b].qe = qe;
        blocks[b].slot = b;
        qb = qe;
      }
    }
    std::mutex mu;
    std::condition_variable cv;
    size_t chained_upto = 0, gapped_upto = 0;
    int device_status = 0;
    std::thread device([&]() {
      for (size_t b = 0; b < nblocks; ++b) {
        {
          std::unique_lock<std::mutex> lock(mu);
          cv.wait(lock, [&] { return chained_upto > b; });
        }
        const double tg = Now();
        const int r = device_status == 0 ? gapped(blocks[b]) : device_status;
        {
          std::lock_guard<std::mutex> lock(mu);
          st->seconds_gapped += Now() - tg;
          if (r != 0) device_status = r;
          gapped_upto = b + 1;
        }
        cv.notify_all();
      }
    });
    for (size_t b = 0; b <= nblocks; ++b) {
      if (b < nblocks) {
        const double tc = Now();
        chain(blocks[b].qb, blocks[b].qe);
        st->seconds_chain += Now() - tc;
        {
          std::lock_guard<std::mutex> lock(mu);
          chained_upto = b + 1;
        }
        cv.notify_all();
      }
      if (b >= 1) {
        {
          std::unique_lock<std::mutex> lock(mu);
          cv.wait(lock, [&] { return gapped_upto >= b; });
        }
        if (device_status == 0) {
          const double ta = Now();
          add(blocks[b - 1]);
          st->seconds_add += Now() - ta;
        }
      }
    }
    device.join();
    GZ_STEP(device_status);
  }

  // BuildResults (on the device where it builds results) and WriteOutput
  t = Now();
  std::vector<std::vector<AlignmentResult> > results(nq);
  for (size_t i = 0; i < nq; ++i) results[i].resize(pre[i].size());
  for (size_t ci = 0; ci < db.chunks.size(); ++ci) {
    const DbChunk &chunk = db.chunks[ci];
    const std::vector<std::pair<uint32_t, uint32_t> > ids = ChunkResultIds(pre, ci);
    if (ids.empty()) continue;
    std::vector<TracebackJob> jobs(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
      const uint32_t qi = ids[k].first;
      const Presearched &pr = pre[qi][ids[k].second];
      jobs[k].query_offset = (uint32_t)coff[qi];
      jobs[k].query_length = q.Length(qi);
      jobs[k].hit_q = pr.hit.q;
      jobs[k].hit_d = pr.hit.d;
    }
    std::vector<TracebackOut> outs;
    const int rc = dev->Tracebacks(ci, jobs, result_cutoff, &outs, st);
    if (rc == PresearchDevice::kTracebackUnsupported) {
      BuildChunkHost(q, chunk, P, result_cutoff, pre, results, ids, threads,
                     &st->sequential_build_chunks);
      continue;
    }
    GZ_STEP(rc);
    bool stale = false;
    std::vector<size_t> host_jobs;
    for (size_t k = 0; k < ids.size(); ++k) {
      if (outs[k].flags & kTracebackStale) stale = true;
      if (outs[k].flags & kTracebackOverflow) host_jobs.push_back(k);
    }
    st->traceback_host_jobs += host_jobs.size();
    if (!stale) {
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
      for (size_t k : host_jobs) {
        const uint32_t qi = ids[k].first, r = ids[k].second;
        if (!BuildOneFresh(q, qi, chunk, pre[qi][r], P, result_cutoff, &results[qi][r])) {
          stale = true;
          break;
        }
      }
    }
    if (stale) {
      ++st->sequential_build_chunks;
      ReplayChunkHost(q, chunk, P, result_cutoff, pre, results, ids);
    }
  }
  SortResults(results);
  dev->EndQueryChunk();
  for (auto &r : results) st->results += r.size();
  st->seconds_build += Now() - t;
  t = Now();
  std::ostringstream os;
  WriteOutput(os, q, db, results);
  output->append(os.str());
  st->seconds_output += Now() - t;
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
