// results.cpp -- see results.h.
#include "results.h"

#include <algorithm>
#include <charconv>
#include <atomic>
#include <climits>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gz_common.h"

namespace gz {

// ---------------------------------------------------------------------------
// chain filter
// ---------------------------------------------------------------------------
namespace {
struct DbOrder {
  bool operator()(const Hit &a, const Hit &b) const { return a.d < b.d; }
};

// ChainFilter::Connected (src/chain_filter.cpp:62-99): does the reverse
// ungapped walk from the hit reach the chain's end before dropping?
bool Connected(const uint8_t *q, const uint8_t *d, const int *sm, uint32_t nl,
               uint32_t chain_end, const Hit &h, int cutoff) {
  if (h.q <= chain_end + 1) return true;
  uint32_t qo = h.q, doff = h.d;
  int score = 0, best = 0, c = -cutoff;
  do {
    --doff;
    --qo;
    if (d[doff] == kDelimiter || q[qo] == kDelimiter || qo <= chain_end) break;
    score += sm[d[doff] * nl + q[qo]];
    if (best < score) {
      best = score;
      c = best - cutoff;
    }
  } while (score > c);
  return qo <= chain_end;
}
}  // namespace

void ChainFilter(const uint8_t *q, uint32_t qlen, int cutoff, const uint8_t *d,
                 const int *sm, uint32_t nl, std::vector<Hit> &hits,
                 std::vector<uint64_t> &chains) {
  std::sort(hits.begin(), hits.end(), DbOrder());
  ChainFilterSorted(q, qlen, cutoff, d, sm, nl, hits, chains);
}

void ChainFilterSorted(const uint8_t *q, uint32_t qlen, int cutoff, const uint8_t *d,
                       const int *sm, uint32_t nl, std::vector<Hit> &hits,
                       std::vector<uint64_t> &chains) {
  // per diagonal slot: (query_end << 32) | diagonal
  chains.assign(qlen, 0);
  size_t n = 0;
  for (size_t i = 0; i < hits.size(); ++i) {
    const Hit h = hits[i];
    const uint32_t diagonal = qlen - 1 + h.d - h.q;
    const uint32_t slot = diagonal % qlen;
    const uint32_t cdiag = (uint32_t)chains[slot];
    const uint32_t cend = (uint32_t)(chains[slot] >> 32);
    if (cdiag != diagonal || !Connected(q, d, sm, nl, cend, h, cutoff)) {
      chains[slot] = ((uint64_t)h.q << 32) | diagonal;
      hits[n++] = h;
    } else {
      chains[slot] = ((uint64_t)h.q << 32) | cdiag;
    }
  }
  hits.resize(n);
}

// ---------------------------------------------------------------------------
// result merge
// ---------------------------------------------------------------------------
namespace {
// AlignerCommon::AlignmentComp (src/aligner_common.cpp:12-56)
int Compare(int as, const Coordinate &a0, const Coordinate &a1, int bs,
            const Coordinate &b0, const Coordinate &b1) {
  if (as != bs) return as > bs ? 1 : -1;
  const uint32_t aq = a1.q - a0.q, bq = b1.q - b0.q;
  if (aq != bq) return aq > bq ? 1 : -1;
  const uint32_t ad = a1.d - a0.d, bd = b1.d - b0.d;
  if (ad != bd) return ad > bd ? 1 : -1;
  return a0.d < b0.d ? 1 : -1;
}
bool Contains(const Coordinate &a0, const Coordinate &a1, const Coordinate &b0,
              const Coordinate &b1) {
  return (a0.q == b0.q && a0.d == b0.d) || (a1.q == b1.q && a1.d == b1.d) ||
         (a0.q <= b0.q && b1.q <= a1.q && a0.d <= b0.d && b1.d <= a1.d);
}
struct ByScore {
  bool operator()(const Presearched &a, const Presearched &b) const {
    return Compare(a.score, a.start, a.end, b.score, b.start, b.end) > 0;
  }
};
struct ResultByScore {
  bool operator()(const AlignmentResult &a, const AlignmentResult &b) const {
    return Compare(a.score, a.start, a.end, b.score, b.start, b.end) > 0;
  }
};
struct PositionLess {
  bool operator()(const AlignmentPosition &a, const AlignmentPosition &b) const {
    return a.position < b.position;
  }
};
}  // namespace

std::vector<AlignmentPosition> &SubjectStarts::Of(uint32_t subject) {
  if (used_ > 16) {
    const auto f = index_.find(subject);
    if (f != index_.end()) return lists_[f->second];
  } else {
    for (size_t i = 0; i < used_; ++i)
      if (subjects_[i] == subject) return lists_[i];
  }
  if (used_ == subjects_.size()) {
    subjects_.push_back(0);
    lists_.emplace_back();
  }
  const size_t slot = used_++;
  subjects_[slot] = subject;
  lists_[slot].clear();
  if (used_ > 16) {
    if (index_.empty())
      for (size_t i = 0; i < used_; ++i) index_[subjects_[i]] = (uint32_t)i;
    else
      index_[subject] = (uint32_t)slot;
  }
  return lists_[slot];
}

void SubjectStarts::Clear() {
  used_ = 0;
  index_.clear();
}

void AddResults(const DbChunk &chunk, int chunk_id, const SearchParameters &p,
                AddResultsScratch &scratch, std::vector<Presearched> &added,
                std::vector<Presearched> &results) {
  std::vector<Presearched> &merged = scratch.merged;
  merged.clear();
  scratch.starts.Clear();
  // The new results are taken in stable score order (the reference's
  // stable_sort), but the merge below stops after max_results kept ones, so
  // only a prefix of that order is produced: a partial sort of indices by
  // (score order, index), whose first k equal the stable sort's first k,
  // extended when the merge runs past them.
  std::vector<uint32_t> &order = scratch.order;
  const size_t na = added.size();
  order.resize(na);
  for (size_t k = 0; k < na; ++k) order[k] = (uint32_t)k;
  auto before = [&](uint32_t x, uint32_t y) {
    // ByScore never calls two results equal; stable_sort keeps the input
    // order of results neither of which comes before the other
    const Presearched &a = added[x], &b = added[y];
    if (Compare(a.score, a.start, a.end, b.score, b.start, b.end) > 0) return true;
    if (Compare(b.score, b.start, b.end, a.score, a.start, a.end) > 0) return false;
    return x < y;
  };
  size_t sorted = 0;  // order[0, sorted) is final
  auto extend = [&](size_t want) {
    want = std::min(na, want);
    if (want <= sorted) return;
    std::partial_sort(order.begin() + sorted, order.begin() + want, order.end(), before);
    sorted = want;
  };
  extend(2 * (size_t)p.max_results + 8);
  size_t ri = 0, ai = 0;
  while (merged.size() < p.max_results && (ai < na || ri < results.size())) {
    if (ai < na && ai == sorted) extend(2 * sorted);
    Presearched a;
    if (ai == na) {
      a = results[ri++];
    } else if (ri == results.size() ||
               Compare(results[ri].score, results[ri].start, results[ri].end,
                       added[order[ai]].score, added[order[ai]].start,
                       added[order[ai]].end) < 0) {
      a = added[order[ai++]];
      a.chunk = chunk_id;
      a.subject = chunk.SubjectOf(a.start.d);
    } else {
      a = results[ri++];
    }
    if (a.chunk == chunk_id) {
      // only this chunk's subjects are entered (subject numbers of results
      // kept from earlier chunks index those chunks)
      std::vector<AlignmentPosition> &sp = scratch.starts.Of(a.subject);
      if (sp.size() < p.max_results_per_subject) {
        AlignmentPosition key = {0, a.start.d};
        auto it = std::lower_bound(sp.begin(), sp.end(), key, PositionLess());
        const auto insert_at = it;
        key.position = a.end.d;
        auto ie = std::lower_bound(it, sp.end(), key, PositionLess());
        bool keep = true;
        for (; it != ie; ++it) {
          const Presearched &r = merged[it->result_id];
          if (Contains(r.start, r.end, a.start, a.end)) {
            keep = false;
            break;
          }
        }
        if (keep) {
          AlignmentPosition ap = {(uint32_t)merged.size(), a.start.q};
          sp.insert(insert_at, ap);
          merged.push_back(a);
        }
      }
    } else {
      merged.push_back(a);
    }
  }
  results.swap(merged);
}

// ---------------------------------------------------------------------------
// traceback re-extension
// ---------------------------------------------------------------------------
namespace {

enum : uint8_t { kSub = 0, kGap0 = 1, kGap1 = 2, kExt0 = 4, kExt1 = 8 };
enum : uint8_t { kEditGap0 = 0, kEditSub = 1, kEditGap1 = 2 };  // EditBlocks order

struct Blocks {
  std::vector<std::pair<uint8_t, int> > b;
  void Add(uint8_t op, int n) {
    if (n == 0) return;
    if (!b.empty() && b.back().first == op) b.back().second += n;
    else b.push_back(std::make_pair(op, n));
  }
};

// GappedExtender with traceback (src/gapped_extender.cpp:70-304,473-531).
// The op store mirrors gap_states_ (:25-68): rows are laid out back to back
// in a byte arena, and the traceback may read bytes outside a row's band;
// such bytes come from neighbouring rows of the same extension.  A read of
// a byte that the current extension did not write would depend on earlier
// extensions run through the same object; those reads are counted so the
// caller can replay the chunk in the reference's order.
class Traceback {
 public:
  uint64_t stale = 0;
  // Bytes of arena 0 that the current extension did not write come from
  // `resolve` when set (it returns false when it cannot tell); `exotic`
  // records an extension that used another arena or a byte past kArena,
  // which the resolver does not model.
  std::function<bool(int64_t idx, uint8_t *v)> resolve;
  bool unresolved = false;
  bool exotic = false;
  int64_t max_put = -1;  // the current extension's highest arena-0 index written

  void Extend(const uint8_t *s0, uint32_t s0len, const uint8_t *s1, bool reversed,
              const int *sm, uint32_t nl, int gap_open, int gap_ext, int cutoff,
              int *best_score, int *best_p0, int *best_p1, Blocks *eb) {
    *best_score = 0;
    const int inc = reversed ? -1 : 1;
    *best_p0 = -inc;
    *best_p1 = -inc;
    eb->b.clear();
    if (s0[0] == kDelimiter || s1[0] == kDelimiter) return;
    uint32_t n0 = 0;
    while (s0[(int)n0 * inc] != kDelimiter) ++n0;
    if (dp_.size() < std::max<size_t>(s0len + 1, n0 + 2)) dp_.resize(std::max<size_t>(s0len + 1, n0 + 2));
    rows_arr_.clear();
    rows_base_.clear();
    rows_off_.clear();
    ++serial_;
    max_put = -1;
    for (auto &a : arenas_) a.used = 0;
    const uint32_t extra = (gap_ext > 0 && cutoff != INT_MAX) ? cutoff / gap_ext + 3 : s0len + 3;
    uint32_t arr = Pick(extra);
    Put(arr, 0, kSub);

    const int gi = gap_open + gap_ext;
    int as = 0, ae = 0, ai = 0, last = 0;
    int score = -gi;
    dp_[0] = {0, -gi};
    int p0 = 0;
    for (ai = 1; s0[p0] != kDelimiter; ++ai, p0 += inc) {
      if (score < -cutoff) break;
      dp_[ai] = {score, score - gi};
      score -= gap_ext;
    }
    if (s0[p0] != kDelimiter) {
      dp_[ai] = {kInitScore, kInitScore};
      ae = ai + 1;
    } else {
      ae = ai;
    }
    int max_score = 0, mp0 = -inc, mp1 = -inc;
    for (int p1 = 0; s1[p1] != kDelimiter; p1 += inc) {
      arr = (gap_ext > 0) ? Pick((uint32_t)(ae - as) + extra) : Pick(extra - (uint32_t)as);
      const int64_t base = (int64_t)arenas_[arr].used + 1;
      const int off = as;
      rows_arr_.push_back(arr);
      rows_base_.push_back(base);
      rows_off_.push_back(off);
      const int64_t zero = base - off;
      const int *row = sm + s1[p1] * nl;
      p0 = as * inc;
      int prev = dp_[as].best;
      dp_[as].best = (max_score - dp_[as].gap > cutoff) ? kInitScore : dp_[as].gap;
      int gap_row = dp_[as].best - gi;
      dp_[as].gap -= gap_ext;
      Put(arr, zero + as, kGap0 + kExt0);
      last = as;
      for (ai = as + 1; ai < ae; ++ai, p0 += inc) {
        int s = prev + row[s0[p0]];
        prev = dp_[ai].best;
        int gap_col = dp_[ai].gap;
        uint8_t op = kSub;
        if (s < gap_col) {
          op = kGap0;
          s = gap_col;
        }
        if (s < gap_row) {
          op = kGap1;
          s = gap_row;
        }
        if (max_score - s > cutoff) {
          if (as + 1 == ai) ++as;
          dp_[ai] = {kInitScore, kInitScore};
          gap_row = kInitScore;
        } else {
          last = ai;
          if (s > max_score) {
            max_score = s;
            mp0 = p0;
            mp1 = p1;
          }
          gap_col -= gap_ext;
          gap_row -= gap_ext;
          if (gap_col < s - gi) {
            dp_[ai].gap = s - gi;
          } else {
            dp_[ai].gap = gap_col;
            op += kExt0;
          }
          if (gap_row < s - gi) gap_row = s - gi; else op += kExt1;
          Put(arr, zero + ai, op);
          dp_[ai].best = s;
        }
      }
      if (as + 1 == ae) break;
      if (last < ae - 1) {
        ae = last + 1;
      } else {
        while (gap_row >= (max_score - cutoff) && s0[p0] != kDelimiter) {
          dp_[ae] = {gap_row, gap_row - gi};
          Put(arr, zero + ae, kGap0 + kExt0);
          gap_row -= gap_ext;
          ++ae;
          p0 += inc;
        }
      }
      if (s0[p0] != kDelimiter) {
        dp_[ae] = {kInitScore, kInitScore};
        ++ae;
      }
      arenas_[arr].used += (uint32_t)(ae - off);
    }
    *best_score = max_score;
    *best_p0 = mp0;
    *best_p1 = mp1;
    // PickOptimalPath
    int q = mp0, d = mp1;
    uint8_t op = kSub;
    while (d * inc >= 0 && q * inc >= 0) {
      const int r = d * inc;
      const uint8_t next = Get(rows_arr_[r], rows_base_[r] + (q * inc) - rows_off_[r] + 1);
      if (op == kGap0) {
        op = next & 3;
        if (next & kExt0) op = kGap0;
      } else if (op == kGap1) {
        op = next & 3;
        if (next & kExt1) op = kGap1;
      } else {
        op = next & 3;
      }
      if (op == kGap0) {
        eb->Add(kEditGap0, 1);
        d -= inc;
      } else if (op == kGap1) {
        eb->Add(kEditGap1, 1);
        q -= inc;
      } else {
        eb->Add(kEditSub, 1);
        d -= inc;
        q -= inc;
      }
    }
    if (d * inc >= 0) eb->Add(kEditGap0, d * inc + 1);
    else if (q * inc >= 0) eb->Add(kEditGap1, q * inc + 1);
    if (inc > 0) std::reverse(eb->b.begin(), eb->b.end());
  }

 private:
  struct Cell {
    int best;
    int gap;
  };
  struct Arena {
    std::vector<uint8_t> a;
    std::vector<uint32_t> stamp;
    uint32_t used = 0;
  };
  static constexpr uint32_t kArena = 1u << 21;

  uint32_t Pick(uint32_t length) {  // GetGapStateIndex
    const uint32_t l = length + length / 3;
    const uint32_t alen = std::max(kArena, l);
    if (arenas_.empty()) {
      arenas_.emplace_back();
      Resize(arenas_[0], alen);
      return 0;
    }
    for (uint32_t i = 0; i < arenas_.size(); ++i) {
      if (l < arenas_[i].a.size() - arenas_[i].used) return i;
      if (arenas_[i].used == 0) {
        Resize(arenas_[i], alen);
        return i;
      }
    }
    arenas_.emplace_back();
    Resize(arenas_.back(), alen);
    return (uint32_t)arenas_.size() - 1;
  }
  static void Resize(Arena &a, uint32_t n) {
    a.a.resize(n);
    a.stamp.resize(n, 0);
  }
  void Put(uint32_t arr, int64_t idx, uint8_t v) {
    arenas_[arr].a[idx] = v;
    arenas_[arr].stamp[idx] = serial_;
    if (arr != 0 || idx >= (int64_t)kArena) exotic = true;
    else if (idx > max_put) max_put = idx;
  }
  uint8_t Get(uint32_t arr, int64_t idx) {
    Arena &a = arenas_[arr];
    if (idx < 0 || idx >= (int64_t)a.a.size()) {
      ++stale;
      if (resolve) {
        // before the arena or past it: 0 in any history while the index is
        // below kArena (arena 0 never holds fewer bytes); otherwise undecided
        if (arr != 0 || idx >= (int64_t)kArena) unresolved = true;
      }
      return 0;
    }
    if (a.stamp[idx] != serial_) {
      ++stale;
      if (resolve) {
        uint8_t v = 0;
        if (arr != 0 || idx >= (int64_t)kArena || !resolve(idx, &v)) unresolved = true;
        return v;
      }
    }
    return a.a[idx];
  }

 public:
  // whether the current (last) extension wrote arena-0 byte idx, and its value
  bool Wrote(int64_t idx, uint8_t *v) const {
    if (arenas_.empty() || idx < 0 || idx >= (int64_t)arenas_[0].a.size()) return false;
    if (arenas_[0].stamp[idx] != serial_) return false;
    *v = arenas_[0].a[idx];
    return true;
  }
  uint32_t Extensions() const { return serial_; }
  bool WroteIndex(int64_t idx) const {
    return !arenas_.empty() && idx >= 0 && idx < (int64_t)arenas_[0].a.size() &&
           arenas_[0].stamp[idx] == serial_;
  }

 private:

  std::vector<Cell> dp_;
  std::vector<Arena> arenas_;
  std::vector<uint32_t> rows_arr_;
  std::vector<int64_t> rows_base_;
  std::vector<int> rows_off_;
  uint32_t serial_ = 0;
};

// BuildResults for one presearched result (aligner_build_results_thread.h
// :97-150 + AlignerCommon::BuildResult, src/aligner_common.h:276-343).
void BuildOneRaw(Traceback &tb, const uint8_t *qs, uint32_t qlen, const uint8_t *ds,
                 const Coordinate hit, const SearchParameters &p, int cutoff,
                 AlignmentResult *out) {
  Blocks rev, fwd;
  int s1, s2, p0, p1;
  tb.Extend(qs + hit.q - 1, hit.q - 1, ds + hit.d - 1, true, p.matrix, p.nl, p.gap_open,
            p.gap_extension, cutoff, &s1, &p0, &p1, &rev);
  Coordinate start = {hit.q - 1 + p0, hit.d - 1 + p1};
  std::reverse(rev.b.begin(), rev.b.end());
  tb.Extend(qs + hit.q, qlen - hit.q, ds + hit.d, false, p.matrix, p.nl, p.gap_open,
            p.gap_extension, cutoff, &s2, &p0, &p1, &fwd);
  Coordinate end = {hit.q + (uint32_t)p0, hit.d + (uint32_t)p1};
  Blocks all;
  for (auto &b : rev.b) all.Add(b.first, b.second);
  for (auto &b : fwd.b) all.Add(b.first, b.second);
  uint32_t qp = start.q, dp = start.d, m = 0, o = 0, e = 0, len = 0;
  uint8_t prev = kEditSub;
  for (auto &b : all.b) {
    for (int k = 0; k < b.second; ++k) {
      const uint8_t op = b.first;
      if (op == kEditSub) {
        if (qs[qp] != ds[dp]) ++m;
        ++qp;
        ++dp;
      } else if (op == kEditGap0) {
        if (prev != kEditGap0) ++o; else ++e;
        ++dp;
      } else {
        if (prev != kEditGap1) ++o; else ++e;
        ++qp;
      }
      prev = op;
      ++len;
    }
  }
  out->score = s1 + s2;
  out->length = len;
  out->mismatches = m;
  out->gap_openings = o;
  out->gap_extensions = e;
  out->hit = hit;
  out->start = start;
  out->end = end;
}

void BuildOne(Traceback &tb, const QuerySet &q, uint32_t qi, const DbChunk &c,
              const Presearched &pr, const SearchParameters &p, int cutoff,
              AlignmentResult *out) {
  BuildOneRaw(tb, q.Sequence(qi), q.Length(qi), c.seq.data(), pr.hit, p, cutoff, out);
  out->chunk = pr.chunk;
  out->subject = pr.subject;
}

}  // namespace

TracebackOut TracebackOneHost(const uint8_t *query, uint32_t query_length,
                              const uint8_t *database, uint32_t hit_q, uint32_t hit_d,
                              const SearchParameters &p, int cutoff) {
  Traceback tb;
  AlignmentResult r;
  const Coordinate hit = {hit_q, hit_d};
  BuildOneRaw(tb, query, query_length, database, hit, p, cutoff, &r);
  TracebackOut o;
  o.score = r.score;
  o.start_q = r.start.q;
  o.start_d = r.start.d;
  o.end_q = r.end.q;
  o.end_d = r.end.d;
  o.length = r.length;
  o.mismatches = r.mismatches;
  o.gap_openings = r.gap_openings;
  o.gap_extensions = r.gap_extensions;
  o.flags = tb.stale ? kTracebackStale : 0;
  return o;
}

bool BuildOneFresh(const QuerySet &q, uint32_t qi, const DbChunk &c, const Presearched &pr,
                   const SearchParameters &p, int cutoff, AlignmentResult *out) {
  Traceback tb;
  BuildOne(tb, q, qi, c, pr, p, cutoff, out);
  return tb.stale == 0;
}

void ReplayChunkHost(const QuerySet &q, const DbChunk &c, const SearchParameters &p, int cutoff,
                     std::vector<std::vector<Presearched> > &pre,
                     std::vector<std::vector<AlignmentResult> > &results,
                     const std::vector<std::pair<uint32_t, uint32_t> > &ids) {
  // one extender object, pairs handed out in blocks of 32 and taken from the
  // back of each block (aligner_build_results_thread.h:97-150, one thread)
  Traceback tb;
  for (size_t b = 0; b < ids.size(); b += 32) {
    const size_t e = std::min(ids.size(), b + 32);
    for (size_t k = e; k-- > b;) {
      const uint32_t qi = ids[k].first, r = ids[k].second;
      BuildOne(tb, q, qi, c, pre[qi][r], p, cutoff, &results[qi][r]);
    }
  }
}

bool ResolveStaleChunk(const QuerySet &q, const DbChunk &c, const SearchParameters &p,
                       int cutoff, std::vector<std::vector<Presearched> > &pre,
                       std::vector<std::vector<AlignmentResult> > &results,
                       const std::vector<std::pair<uint32_t, uint32_t> > &ids,
                       const std::vector<size_t> &flagged) {
  constexpr uint32_t kMaxLookback = 256;
  const size_t n = ids.size();
  // ReplayChunkHost's order: blocks of 32, each from its back; extension
  // 2t is the reverse side of the t-th result in that order, 2t + 1 its
  // forward side
  std::vector<size_t> order;
  std::vector<size_t> place(n);
  order.reserve(n);
  for (size_t b = 0; b < n; b += 32) {
    const size_t e = std::min(n, b + 32);
    for (size_t k = e; k-- > b;) {
      place[k] = order.size();
      order.push_back(k);
    }
  }
  // one extension on its own: the arena-0 bytes it writes (false when it
  // uses more than arena 0's first kArena bytes)
  struct Writes {
    std::vector<uint8_t> bytes;
    std::vector<uint8_t> wrote;
  };
  std::unordered_map<size_t, Writes> cache;
  auto writes_of = [&](size_t x, const Writes **out) -> bool {
    auto it = cache.find(x);
    if (it != cache.end()) {
      *out = &it->second;
      return true;
    }
    const std::pair<uint32_t, uint32_t> id = ids[order[x / 2]];
    const Presearched &pr = pre[id.first][id.second];
    const uint8_t *qs = q.Sequence(id.first);
    const uint32_t qlen = q.Length(id.first);
    const uint8_t *ds = c.seq.data();
    Traceback tb;
    Blocks eb;
    int sc, p0, p1;
    if (x % 2 == 0)
      tb.Extend(qs + pr.hit.q - 1, pr.hit.q - 1, ds + pr.hit.d - 1, true, p.matrix, p.nl,
                p.gap_open, p.gap_extension, cutoff, &sc, &p0, &p1, &eb);
    else
      tb.Extend(qs + pr.hit.q, qlen - pr.hit.q, ds + pr.hit.d, false, p.matrix, p.nl,
                p.gap_open, p.gap_extension, cutoff, &sc, &p0, &p1, &eb);
    if (tb.exotic) return false;
    Writes w;
    const int64_t top = tb.max_put + 1;
    w.bytes.resize((size_t)std::max<int64_t>(top, 0));
    w.wrote.resize(w.bytes.size());
    for (int64_t i = 0; i < top; ++i) {
      uint8_t v = 0;
      if (tb.Wrote(i, &v)) {
        w.bytes[(size_t)i] = v;
        w.wrote[(size_t)i] = 1;
      }
    }
    if (cache.size() > 64) cache.clear();
    *out = &cache.emplace(x, std::move(w)).first->second;
    return true;
  };
  for (size_t k : flagged) {
    const std::pair<uint32_t, uint32_t> id = ids[k];
    const size_t t = place[k];
    Traceback tb;
    // byte idx as the extensions before the current one left it
    tb.resolve = [&](int64_t idx, uint8_t *v) -> bool {
      const size_t x = 2 * t + (tb.Extensions() - 1);
      for (size_t back = 1; back <= x; ++back) {
        if (back > kMaxLookback) return false;
        const Writes *w = nullptr;
        if (!writes_of(x - back, &w)) return false;
        if ((size_t)idx < w->wrote.size() && w->wrote[(size_t)idx]) {
          *v = w->bytes[(size_t)idx];
          return true;
        }
      }
      *v = 0;  // no earlier extension wrote it: the zeroed arena
      return true;
    };
    AlignmentResult r;
    BuildOne(tb, q, id.first, c, pre[id.first][id.second], p, cutoff, &r);
    if (tb.unresolved || tb.exotic) return false;
    results[id.first][id.second] = r;
  }
  return true;
}

void BuildChunkHost(const QuerySet &q, const DbChunk &c, const SearchParameters &p, int cutoff,
                    std::vector<std::vector<Presearched> > &pre,
                    std::vector<std::vector<AlignmentResult> > &results,
                    const std::vector<std::pair<uint32_t, uint32_t> > &ids, int threads,
                    uint64_t *sequential_chunks) {
  if (ids.empty()) return;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  threads = std::max(1, threads);
  std::atomic<size_t> next(0);
  std::atomic<uint64_t> stale(0);
  const int T = std::max(1, std::min<int>(threads, (int)(ids.size() / 64) + 1));
  auto work = [&]() {
    Traceback tb;
    for (;;) {
      const size_t b = next.fetch_add(256);
      if (b >= ids.size()) break;
      for (size_t k = b; k < std::min(ids.size(), b + 256); ++k) {
        const uint32_t qi = ids[k].first, r = ids[k].second;
        BuildOne(tb, q, qi, c, pre[qi][r], p, cutoff, &results[qi][r]);
      }
    }
    stale += tb.stale;
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (stale.load() != 0) {
    // history-dependent bytes were read: replay the chunk exactly as the
    // reference's one thread does
    if (sequential_chunks) ++*sequential_chunks;
    ReplayChunkHost(q, c, p, cutoff, pre, results, ids);
  }
}

std::vector<std::pair<uint32_t, uint32_t> > ChunkResultIds(
    const std::vector<std::vector<Presearched> > &pre, size_t ci) {
  return ChunkResultIds(pre, ci, 0, pre.size());
}

std::vector<std::pair<uint32_t, uint32_t> > ChunkResultIds(
    const std::vector<std::vector<Presearched> > &pre, size_t ci, size_t qb, size_t qe) {
  std::vector<std::pair<uint32_t, uint32_t> > ids;
  for (uint32_t qi = (uint32_t)qb; qi < qe; ++qi)
    for (uint32_t r = 0; r < pre[qi].size(); ++r)
      if (pre[qi][r].chunk == (int)ci) ids.push_back(std::make_pair(qi, r));
  return ids;
}

void SortResults(std::vector<std::vector<AlignmentResult> > &results) {
  SortResults(results, 0, results.size());
}

void SortResults(std::vector<std::vector<AlignmentResult> > &results, size_t qb, size_t qe) {
  for (size_t i = qb; i < qe; ++i) std::sort(results[i].begin(), results[i].end(), ResultByScore());
}

void BuildResults(const QuerySet &q, const Database &db, const SearchParameters &p,
                  int cutoff, std::vector<std::vector<Presearched> > &pre,
                  std::vector<std::vector<AlignmentResult> > &results, int threads,
                  uint64_t *sequential_chunks) {
  const size_t nq = q.size();
  results.assign(nq, std::vector<AlignmentResult>());
  for (size_t i = 0; i < nq; ++i) results[i].resize(pre[i].size());
  for (size_t ci = 0; ci < db.chunks.size(); ++ci)
    BuildChunkHost(q, db.chunks[ci], p, cutoff, pre, results, ChunkResultIds(pre, ci), threads,
                   sequential_chunks);
  SortResults(results);
}

// ---------------------------------------------------------------------------
// output
// ---------------------------------------------------------------------------
void WriteOutput(std::ostream &os, const QuerySet &q, const Database &db,
                 const std::vector<std::vector<AlignmentResult> > &results) {
  std::string text;
  WriteOutputText(&text, q, db, results, 0, q.size(), 0);
  os << text;
}

namespace {
// 10^k for k in [kPow10Low, kPow10Low + kPow10Count), each the double nearest
// the decimal value (strtod rounds correctly)
constexpr int kPow10Low = -296, kPow10Count = 604;
const double *Pow10Table() {
  static const std::vector<double> table = [] {
    std::vector<double> t(kPow10Count);
    for (int i = 0; i < kPow10Count; ++i) {
      const std::string text = "1e" + std::to_string(kPow10Low + i);
      t[i] = std::strtod(text.c_str(), nullptr);
    }
    return t;
  }();
  return table.data();
}
}  // namespace

char *FormatG6(double v, char *p, char *end) {
  const double a = std::fabs(v);
  // the fast path covers 1e-300 <= |v| < 1e300; zero, signed zero, the
  // extremes and non-finite values are left to the library
  if (!(a >= 1e-300 && a < 1e300) || end - p < 16)
    return std::to_chars(p, end, v, std::chars_format::general, 6).ptr;
  // decimal exponent: a in [2^(e-1), 2^e) puts log10(a) within 0.302 above
  // (e - 1) log10(2), so x is floor((e - 1) log10(2)) or one more
  // (floor(n log10 2) = (n * 78913) >> 18 for |n| < 1200)
  uint64_t bits;
  std::memcpy(&bits, &a, sizeof(bits));
  const int e = (int)((bits >> 52) & 0x7ff) - 1022;
  int x = ((e - 1) * 78913) >> 18;
  const double *pow10 = Pow10Table();
  uint32_t n = 0;
  for (int tries = 0;; ++tries) {
    // a * 10^(5 - x) below 10^7, with a relative error of at most 2^-52
    // (one rounding of the power, one of the product): within 3e-9 of the
    // exact product, so a fraction further than 1e-6 from one half rounds
    // the way the exact product does; nearer ties are left to the library
    const int k = 5 - x - kPow10Low;
    if (tries == 2 || k < 0 || k >= kPow10Count)
      return std::to_chars(p, end, v, std::chars_format::general, 6).ptr;
    const double scaled = a * pow10[k];
    const uint64_t ip = (uint64_t)(int64_t)scaled;
    const double f = scaled - (double)ip;  // exact: scaled < 2^53
    if (f > 0.5 - 1e-6 && f < 0.5 + 1e-6)
      return std::to_chars(p, end, v, std::chars_format::general, 6).ptr;
    const uint64_t r = f < 0.5 ? ip : ip + 1;
    if (r >= 1000000) {
      ++x;
    } else if (r < 100000) {
      --x;
    } else {
      n = (uint32_t)r;
      break;
    }
  }
  char d[6];
  for (int i = 5; i >= 0; --i) {
    d[i] = (char)('0' + n % 10);
    n /= 10;
  }
  int nd = 6;  // significant digits left once trailing zeros go
  while (nd > 1 && d[nd - 1] == '0') --nd;
  if (v < 0) *p++ = '-';
  if (x < -4 || x >= 6) {
    *p++ = d[0];
    if (nd > 1) {
      *p++ = '.';
      for (int i = 1; i < nd; ++i) *p++ = d[i];
    }
    *p++ = 'e';
    *p++ = x < 0 ? '-' : '+';
    const int ex = x < 0 ? -x : x;  // at least two digits
    if (ex >= 100) *p++ = (char)('0' + ex / 100);
    *p++ = (char)('0' + ex / 10 % 10);
    *p++ = (char)('0' + ex % 10);
  } else if (x >= 0) {
    for (int i = 0; i <= x; ++i) *p++ = d[i];
    if (nd > x + 1) {
      *p++ = '.';
      for (int i = x + 1; i < nd; ++i) *p++ = d[i];
    }
  } else {
    *p++ = '0';
    *p++ = '.';
    for (int i = 0; i < -x - 1; ++i) *p++ = '0';
    for (int i = 0; i < nd; ++i) *p++ = d[i];
  }
  return p;
}

void WriteOutputText(std::string *out, const QuerySet &q, const Database &db,
                     const std::vector<std::vector<AlignmentResult> > &results, size_t qb,
                     size_t qe, int threads) {
  const KarlinParameters gk = Statistics::GappedBlosum62();
  const double alpha = 1.9, beta = -30;
  const size_t nq = qe;
  const size_t block = 256;
  const size_t nblocks = (qe - qb + block - 1) / block;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  const int T = std::max(1, std::min<int>(threads, (int)nblocks));
  // each thread formats its blocks into one buffer of its own, reserved once
  // (a line holds the query name and about 50 bytes of numbers; 96 more
  // bytes cover a subject name of 40), and the blocks are then copied out
  // in query order
  size_t estimate = 0;
  for (size_t qi = qb; qi < qe; ++qi)
    estimate += results[qi].size() * (q.names[qi].size() + 96);
  struct Piece {
    int buffer;
    size_t begin, end;
  };
  std::vector<Piece> pieces(nblocks);
  std::vector<std::string> buffers(T);
  std::atomic<size_t> next(0);
  auto work = [&](int t) {
    std::string &s = buffers[t];
    s.reserve(estimate / T + estimate / (4 * T) + 4096);
    for (;;) {
      const size_t b = next.fetch_add(1);
      if (b >= nblocks) break;
      pieces[b].buffer = t;
      pieces[b].begin = s.size();
      char num[64];
      // the fields as std::ostream writes them: integers in decimal,
      // floating-point values as printf's %g (6 significant digits)
      auto put_u = [&](uint64_t v) {
        const auto r = std::to_chars(num, num + sizeof(num), v);
        s.append(num, r.ptr);
      };
      auto put_g = [&](double v) { s.append(num, FormatG6(v, num, num + sizeof(num))); };
      for (size_t qi = qb + b * block; qi < std::min(nq, qb + (b + 1) * block); ++qi) {
        if (results[qi].empty()) continue;
        const uint64_t space = Statistics::SearchSpace(q.RealLength(qi), db.total_length,
                                                       db.total_sequences, gk, alpha, beta);
        for (const AlignmentResult &r : results[qi]) {
          const DbChunk &c = db.chunks[r.chunk];
          const uint32_t base = c.offsets[r.subject] - 1;
          s += q.names[qi];
          s += '\t';
          s += c.names[r.subject];
          s += '\t';
          put_g((1.0 - (static_cast<float>(r.mismatches + r.gap_openings + r.gap_extensions) /
                        static_cast<float>(r.length))) *
                100);
          s += '\t';
          put_u(r.length);
          s += '\t';
          put_u(r.mismatches);
          s += '\t';
          put_u(r.gap_openings);
          s += '\t';
          put_u(q.RealStart(qi, r.start.q));
          s += '\t';
          put_u(q.RealEnd(qi, r.end.q));
          s += '\t';
          put_u((uint32_t)(r.start.d - base));
          s += '\t';
          put_u((uint32_t)(r.end.d - base));
          s += '\t';
          put_g(Statistics::Nominal2EValue(r.score, space, gk));
          s += '\t';
          put_g((double)Statistics::Nominal2Normalized(r.score, gk));
          s += '\n';
        }
      }
      pieces[b].end = s.size();
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  size_t bytes = out->size();
  for (const Piece &pc : pieces) bytes += pc.end - pc.begin;
  if (out->capacity() < bytes) out->reserve(std::max(bytes, 2 * out->capacity()));
  for (const Piece &pc : pieces) out->append(buffers[pc.buffer], pc.begin, pc.end - pc.begin);
}

}  // namespace gz
