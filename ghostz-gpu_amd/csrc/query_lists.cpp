// query_lists.cpp -- see query_lists.h.
#include "query_lists.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "gz_common.h"

namespace gz {

void ConcatenatedOffsets(const QuerySet &q, std::vector<uint64_t> *offsets,
                         std::vector<uint8_t> *concat) {
  const size_t n = q.size();
  offsets->resize(n);
  uint64_t total = 1;
  for (size_t i = 0; i < n; ++i) total += q.Length(i) - 1;
  concat->resize(total);
  uint64_t off = 0;
  for (size_t i = 0; i < n; ++i) {
    (*offsets)[i] = off;
    memcpy(concat->data() + off, q.Sequence(i), q.Length(i) - 1);
    off += q.Length(i) - 1;
  }
  (*concat)[off] = kDelimiter;
}

namespace {

// f(i) for i in [0, n) on up to `threads` threads, in blocks of `grain`
template <class F>
void ParallelFor(size_t n, int threads, size_t grain, F f) {
  const int T = std::max(1, std::min<int>(threads, (int)((n + grain - 1) / std::max<size_t>(grain, 1))));
  std::atomic<size_t> next(0);
  auto work = [&]() {
    for (;;) {
      const size_t b = next.fetch_add(grain);
      if (b >= n) break;
      const size_t e = std::min(n, b + grain);
      for (size_t i = b; i < e; ++i) f(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
}

struct Entry {
  uint32_t hash;
  uint32_t pos;   // concatenated
  uint32_t qid;
  uint32_t check;
};

void StableSortByHash(std::vector<Entry> &a, int threads) {
  const size_t n = a.size();
  if (n < 2) return;
  std::vector<Entry> tmp(n);
  const int T = std::max(1, std::min<int>(threads, (int)(n / 65536) + 1));
  const size_t per = (n + T - 1) / T;
  std::vector<std::vector<size_t> > hist(T, std::vector<size_t>(65536));
  for (int pass = 0; pass < 2; ++pass) {
    const int shift = pass * 16;
    std::vector<Entry> &src = pass == 0 ? a : tmp;
    std::vector<Entry> &dst = pass == 0 ? tmp : a;
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
      pool.emplace_back([&, t]() {
        std::fill(hist[t].begin(), hist[t].end(), 0);
        const size_t b = t * per, e = std::min(n, b + per);
        for (size_t i = b; i < e; ++i) ++hist[t][(src[i].hash >> shift) & 0xffff];
      });
    for (auto &p : pool) p.join();
    pool.clear();
    size_t sum = 0;
    for (int d = 0; d < 65536; ++d)
      for (int t = 0; t < T; ++t) {
        const size_t c = hist[t][d];
        hist[t][d] = sum;
        sum += c;
      }
    for (int t = 0; t < T; ++t)
      pool.emplace_back([&, t]() {
        const size_t b = t * per, e = std::min(n, b + per);
        for (size_t i = b; i < e; ++i) dst[hist[t][(src[i].hash >> shift) & 0xffff]++] = src[i];
      });
    for (auto &p : pool) p.join();
  }
}

}  // namespace

void BuildQueryLists(const QuerySet &q, const std::vector<uint64_t> &coff,
                     const Database &db, int threads, QueryLists *out) {
  const size_t n = q.size();
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  threads = std::max(1, threads);
  const uint32_t L = db.subsequence_length;
  // appearance order of queries: blocks of 64, reversed within each block
  std::vector<uint32_t> order(n);
  for (size_t b = 0; b < n; b += 64) {
    const size_t e = std::min(n, b + 64);
    for (size_t k = 0; k < e - b; ++k) order[b + k] = (uint32_t)(e - 1 - k);
  }
  // hash every query position (SeedSearcherQueryParameters::Build :45-100)
  std::vector<std::vector<Entry> > per(n);
  std::atomic<size_t> next(0);
  auto work = [&]() {
    for (;;) {
      const size_t b = next.fetch_add(256);
      if (b >= n) break;
      for (size_t qi = b; qi < std::min(n, b + 256); ++qi) {
        const uint8_t *s = q.Sequence(qi);
        const uint32_t len = q.Length(qi);
        std::vector<Entry> &v = per[qi];
        v.clear();
        uint32_t off = 0;
        while (off < len) {
          uint32_t e = off;
          while (e < len && s[e] != kDelimiter) ++e;
          if (e - off >= L) {
            const uint32_t cs = off + L / 2, ce = e - L / 2;
            for (uint32_t i = off; i < e; ++i) {
              uint32_t h, hl;
              if (db.hash.Hash(s + i, &h, &hl)) {
                Entry en;
                en.hash = h;
                en.pos = (uint32_t)(coff[qi] + i);
                en.qid = (uint32_t)qi;
                en.check = (cs <= i && i < ce) ? 1 : 0;
                v.push_back(en);
              }
            }
          }
          off = e + 1;
        }
      }
    }
  };
  {
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
  }
  // appearance order: queries in `order`, entries of a query in position order
  std::vector<size_t> qoff(n + 1, 0);
  for (size_t k = 0; k < n; ++k) qoff[k + 1] = qoff[k] + per[order[k]].size();
  const size_t total = qoff[n];
  std::vector<Entry> all(total);
  // the appearance index rides through the stable sort in the check word's
  // upper bits
  ParallelFor(n, threads, 64, [&](size_t k) {
    std::vector<Entry> &v = per[order[k]];
    Entry *dst = all.data() + qoff[k];
    for (size_t j = 0; j < v.size(); ++j) {
      dst[j] = v[j];
      dst[j].check |= (uint32_t)((qoff[k] + j) << 1);
    }
    std::vector<Entry>().swap(v);
  });
  StableSortByHash(all, threads);
  // groups of equal hashes (sorted by hash, appearance order inside)
  std::vector<uint32_t> gstart;
  {
    const size_t T = (size_t)std::max(1, threads);
    const size_t span = (total + T - 1) / std::max<size_t>(T, 1);
    std::vector<std::vector<uint32_t> > part(T);
    ParallelFor(T, threads, 1, [&](size_t t) {
      const size_t b = t * span, e = std::min(total, b + span);
      for (size_t i = b; i < e; ++i)
        if (i == 0 || all[i].hash != all[i - 1].hash) part[t].push_back((uint32_t)i);
    });
    for (auto &v : part) gstart.insert(gstart.end(), v.begin(), v.end());
  }
  const size_t nl = gstart.size();
  gstart.push_back((uint32_t)total);
  // list ids in first-appearance order without a sort: the first entry of a
  // group is its earliest appearance; walking appearance indexes in order
  // meets the groups in that order
  std::vector<uint32_t> group_at(total, 0xffffffffu);
  ParallelFor(nl, threads, 4096, [&](size_t g) {
    group_at[all[gstart[g]].check >> 1] = (uint32_t)g;
  });
  std::vector<uint32_t> by_appearance;
  by_appearance.reserve(nl);
  for (size_t i = 0; i < total; ++i)
    if (group_at[i] != 0xffffffffu) by_appearance.push_back(group_at[i]);
  std::vector<uint32_t>().swap(group_at);
  // list id (first-appearance order) -> rank (presearch visiting order)
  std::vector<uint32_t> group_of_rank(nl);
  for (size_t b = 0; b < nl; b += 256) {
    const size_t e = std::min(nl, b + 256);
    for (size_t k = 0; k < e - b; ++k) group_of_rank[b + k] = by_appearance[e - 1 - k];
  }
  out->hash.resize(nl);
  out->entry_off.resize(nl + 1);
  out->entry_pos.resize(total);
  out->entry_qid.resize(total);
  out->entry_check.resize(total);
  size_t o = 0;
  for (size_t r = 0; r < nl; ++r) {
    const uint32_t g = group_of_rank[r];
    out->hash[r] = all[gstart[g]].hash;
    out->entry_off[r] = (uint32_t)o;
    o += gstart[g + 1] - gstart[g];
  }
  out->entry_off[nl] = (uint32_t)o;
  ParallelFor(nl, threads, 1024, [&](size_t r) {
    const uint32_t g = group_of_rank[r];
    size_t oo = out->entry_off[r];
    for (uint32_t i = gstart[g]; i < gstart[g + 1]; ++i, ++oo) {
      out->entry_pos[oo] = all[i].pos;
      out->entry_qid[oo] = all[i].qid;
      out->entry_check[oo] = (uint8_t)(all[i].check & 1);
    }
  });
}

}  // namespace gz
