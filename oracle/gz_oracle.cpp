/*
 * gz_oracle.cpp -- CPU oracle for the GHOSTZ-GPU `aln` search path.
 *
 * TEST INFRASTRUCTURE ONLY.  A plain, single-threaded restatement of the
 * reference CPU aligner (`ghostz-gpu aln -g 0 -a 1`), used as the parity
 * checker of the HIP product path.  Nothing in ghostz-gpu_amd/ links or loads
 * it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Pinning (what keeps this restatement honest):
 *   - extenders, distance, hash, statistics: bit-exact against the
 *     reference's own sources compiled into oracle/_ref (see ref_driver.cpp)
 *     on randomised inputs, plus the known answers of the reference tests;
 *   - chain filter: reference tests test/chain_filter_test.cpp:11-121 plus a
 *     brute-force check of the X-drop walk;
 *   - the whole search (translation, hash lists and their visiting order,
 *     chain filter, AddResults, BuildResults, output fields): the reference's
 *     own end-to-end example, README.md:88-97 on data/*.fasta
 *     (tests/test_readme_example.py);
 *   - AddResults alone: known answers worked by hand from
 *     src/aligner_common.h:147-274 for every rule, including its containment
 *     quirk (tests/test_add_results.py);
 *   - seed enumeration order and BuildResults' arena order: the reference
 *     does not compile standalone (Boost), so they stay pinned only by
 *     reading (parity unpinned beyond the README example).
 * Parity hazard decisions (recorded once):
 *   - presearch `end` (uninitialised in aligner_presearch_thread.cpp:322,
 *     363-364) is defined as hit + forward best offset, as the GPU path does
 *     (aligner_gpu_presearch_thread.cpp:969-973); gzo_set_end_reading(1)
 *     switches to the other reading (zero at the declaration, accumulated
 *     over the hits of one Extend call) to count what that would change;
 *   - order = `-a 1` order.
 */
#include "gz_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

const int kInitScore = -(1 << 10);  // gapped_extender.h:40

// ---------------------------------------------------------------------------
// UngappedExtender (src/ungapped_extender.h:34-143)
// ---------------------------------------------------------------------------
int UngappedScoreOnly(const uint8_t *s0, const uint8_t *s1, uint8_t delim,
                      bool reversed, const int *sm, uint32_t nl, int cutoff) {
  const int inc = reversed ? -1 : 1;
  const int threshold = -cutoff;
  int pos = 0;
  int best = 0;
  if (s1[pos] != delim && s0[pos] != delim) {
    int score = 0;
    do {
      if (s1[pos] == delim || s0[pos] == delim) break;
      score += sm[s1[pos] * nl + s0[pos]];
      pos += inc;
      if (score > 0) {
        do {
          best += score;
          if (s1[pos] == delim || s0[pos] == delim) break;
          score = sm[s1[pos] * nl + s0[pos]];
          pos += inc;
        } while (score > 0);
      }
    } while (score > threshold);
  }
  return best;
}

int UngappedWithPositions(const uint8_t *s0, const uint8_t *s1, uint8_t delim,
                          bool reversed, const int *sm, uint32_t nl,
                          int cutoff, int *p0, int *p1) {
  const int inc = reversed ? -1 : 1;
  const int threshold = -cutoff;
  int pos = 0;
  int best_pos = 0;
  int best = 0;
  if (s1[pos] != delim && s0[pos] != delim) {
    int score = 0;
    do {
      if (s1[pos] == delim || s0[pos] == delim) break;
      score += sm[s1[pos] * nl + s0[pos]];
      pos += inc;
      if (score > 0) {
        do {
          best += score;
          best_pos = pos;
          if (s1[pos] == delim || s0[pos] == delim) break;
          score = sm[s1[pos] * nl + s0[pos]];
          pos += inc;
        } while (score > 0);
      }
    } while (score > threshold);
  }
  best_pos -= inc;
  *p0 = best_pos;
  *p1 = best_pos;
  return best;
}

// SeedSearcher::UngappedExtend (src/seed_searcher.h:133-158)
int UngappedWithTrigger(const uint8_t *q, uint32_t qpos, const uint8_t *d,
                        uint32_t dpos, uint8_t delim, const int *sm,
                        uint32_t nl, int cutoff, int trigger) {
  int sum = UngappedScoreOnly(q + qpos - 1, d + dpos - 1, delim, true, sm, nl,
                              cutoff);
  if (sum <= trigger) {
    sum += UngappedScoreOnly(q + qpos, d + dpos, delim, false, sm, nl, cutoff);
  }
  return sum;
}

// DistanceCalculator::CalculateDistance (src/distance_calculator.cpp:35-54)
int Distance(const uint8_t *q, uint32_t qc, const uint8_t *d, uint32_t dc,
             uint32_t L, const uint8_t *map) {
  const uint32_t half = L / 2;
  const uint8_t *a = q + qc - half;
  const uint8_t *b = d + dc - half;
  char count = 0;  // DistanceCalculator::Distance is char
  // map holds the 25 letter codes; a delimiter (or any code past them)
  // differs from every letter and equals only itself, and is never read
  // from map (which has no entry for it)
  auto reduced = [&](uint8_t c) { return c < 25 ? (int)map[c] : 256 + c; };
  for (uint32_t i = 0; i < L; ++i) count += (reduced(a[i]) != reduced(b[i])) ? 1 : 0;
  return count;
}

// ---------------------------------------------------------------------------
// GappedExtender (src/gapped_extender.cpp:70-531)
// ---------------------------------------------------------------------------
enum { kOpSub = 0, kOpGap0 = 1, kOpGap1 = 2, kOpExt0 = 4, kOpExt1 = 8 };
// EditBlocks::EditOpType order (src/edit_blocks.h:14-18)
enum { kEditGap0 = 0, kEditSub = 1, kEditGap1 = 2 };

struct EditBlocks {  // src/edit_blocks.cpp
  std::vector<std::pair<int, int> > blocks;
  void Add(int op, int len) {
    if (len == 0) return;
    if (!blocks.empty() && blocks.back().first == op)
      blocks.back().second += len;
    else
      blocks.push_back(std::make_pair(op, len));
  }
  void Add(const EditBlocks &o) {
    for (size_t i = 0; i < o.blocks.size(); ++i)
      Add(o.blocks[i].first, o.blocks[i].second);
  }
  void Reverse() { std::reverse(blocks.begin(), blocks.end()); }
  std::vector<int> ToVector() const {
    std::vector<int> v;
    for (size_t i = 0; i < blocks.size(); ++i)
      for (int k = 0; k < blocks[i].second; ++k) v.push_back(blocks[i].first);
    return v;
  }
};

struct Cell {
  int best;
  int best_gap;
};

// The traceback op storage of GappedExtender (gap_states_,
// src/gapped_extender.cpp:25-68): a list of byte arenas with a `used` mark;
// each DP row's ops are placed at arena[used + 1 ...] and `used` advances by
// the row's width.  PickOptimalPath may read bytes outside a row's band; they
// then come from neighbouring rows, or from earlier extensions run through the
// same extender (the arenas are cleared by resetting `used` only).  The arena
// is therefore restated exactly and kept per extender object.
struct GapArena {
  struct State {
    std::vector<uint8_t> a;
    uint32_t used;
    std::vector<uint64_t> stamp;  // extension serial that last wrote a byte
  };
  static const uint32_t kMaxStateArrayLength = 1u << 21;
  std::vector<State> st;
  uint64_t serial = 0;
  uint64_t stale_reads = 0;
  void Clear() {
    for (size_t i = 0; i < st.size(); ++i) st[i].used = 0;
  }
  void Resize(State &s, uint32_t n) {
    s.a.resize(n);
    s.stamp.resize(n, 0);
  }
  uint32_t Index(uint32_t length) {  // GetGapStateIndex
    uint32_t l = length + length / 3;
    uint32_t alen = std::max(kMaxStateArrayLength, l);
    if (st.empty()) {
      st.push_back(State());
      st[0].used = 0;
      Resize(st[0], alen);
      return 0;
    }
    uint32_t r = (uint32_t)st.size();
    for (uint32_t i = 0; i < st.size(); ++i) {
      if (l < (st[i].a.size() - st[i].used)) {
        r = i;
        break;
      } else if (st[i].used == 0) {
        Resize(st[i], alen);
        r = i;
        break;
      }
    }
    if (r == st.size()) {
      st.push_back(State());
      st.back().used = 0;
      Resize(st.back(), alen);
    }
    return r;
  }
  void Write(uint32_t arr, int64_t idx, uint8_t v) {
    st[arr].a[idx] = v;
    st[arr].stamp[idx] = serial;
  }
  uint8_t Read(uint32_t arr, int64_t idx) {
    if (idx < 0 || idx >= (int64_t)st[arr].a.size()) {
      ++stale_reads;  // out of the arena: undefined in the reference
      return 0;
    }
    if (st[arr].stamp[idx] != serial) ++stale_reads;
    return st[arr].a[idx];
  }
};

// One-sided affine X-drop.  kTrace selects the traceback variant
// (:70-304 + PickOptimalPath :473-531); otherwise score-only (:306-471).
template <bool kTrace>
bool GappedOneSide(const uint8_t *s0, uint32_t s0len, const uint8_t *s1,
                   uint8_t delim, bool reversed, const int *sm, uint32_t nl,
                   int gap_open, int gap_ext, int cutoff, int *best_score,
                   int *best_p0, int *best_p1, EditBlocks *eb,
                   uint64_t *cells, GapArena *arena) {
  *best_score = 0;
  const int inc = reversed ? -1 : 1;
  *best_p0 = -inc;
  *best_p1 = -inc;
  if (kTrace) eb->blocks.clear();
  if (cells) *cells = 0;
  if (s0[0] == delim || s1[0] == delim) return false;

  // letters available in seq0 before its delimiter
  uint32_t n0 = 0;
  while (s0[(int)n0 * inc] != delim) ++n0;
  std::vector<Cell> dp(std::max<size_t>(s0len + 1, n0 + 2));
  // traceback storage: per seq1 row, the arena, the arena index of the
  // row's first op and the row's starting column (edit_script /
  // edit_script_offset of the reference)
  std::vector<uint32_t> row_arr;
  std::vector<int64_t> row_base;
  std::vector<int> row_offset;
  uint32_t extra = 0;
  uint32_t arr = 0;
  if (kTrace) {
    extra = (gap_ext > 0 && cutoff != INT_MAX) ? cutoff / gap_ext + 3
                                               : s0len + 3;
    ++arena->serial;
    arena->Clear();
    arr = arena->Index(extra);
    arena->Write(arr, 0, kOpSub);
  }

  int array_start = 0, array_end = 0, ai = 0, last = 0;
  const int gap_init = gap_open + gap_ext;
  int score = -gap_init;
  dp[0].best = 0;
  dp[0].best_gap = -gap_init;
  int p0 = 0;
  for (ai = 1; s0[p0] != delim; ++ai, p0 += inc) {
    if (score < -cutoff) break;
    dp[ai].best = score;
    dp[ai].best_gap = score - gap_init;
    score -= gap_ext;
  }
  if (s0[p0] != delim) {
    dp[ai].best = kInitScore;
    dp[ai].best_gap = kInitScore;
    array_end = ai + 1;
  } else {
    array_end = ai;
  }
  int max_score = 0;
  int max_p0 = -inc, max_p1 = -inc;
  int prev = 0, gap_row = 0, gap_col = 0;
  uint64_t ncells = 0;
  int p1;
  for (p1 = 0; s1[p1] != delim; p1 += inc) {
    const int row_off = array_start;
    int64_t row_zero = 0;  // arena index of column 0 of this row
    if (kTrace) {
      if (gap_ext > 0) {
        arr = arena->Index((uint32_t)(array_end - array_start) + extra);
      } else {
        arr = arena->Index(extra - (uint32_t)array_start);
      }
      const int64_t base = (int64_t)arena->st[arr].used + 1;
      row_arr.push_back(arr);
      row_base.push_back(base);
      row_offset.push_back(array_start);
      row_zero = base - row_off;
    }
    const int *mrow = sm + s1[p1] * nl;
    p0 = array_start * inc;
    prev = dp[array_start].best;
    if (max_score - dp[array_start].best_gap > cutoff) {
      dp[array_start].best = kInitScore;
    } else {
      dp[array_start].best = dp[array_start].best_gap;
    }
    gap_row = dp[array_start].best - gap_init;
    dp[array_start].best_gap -= gap_ext;
    if (kTrace) arena->Write(arr, row_zero + array_start, kOpGap0 + kOpExt0);
    last = array_start;
    score = kInitScore;
    for (ai = array_start + 1; ai < array_end; ++ai, p0 += inc) {
      ++ncells;
      score = prev + mrow[s0[p0]];
      prev = dp[ai].best;
      gap_col = dp[ai].best_gap;
      uint8_t op = kOpSub;
      if (score < gap_col) {
        op = kOpGap0;
        score = gap_col;
      }
      if (score < gap_row) {
        op = kOpGap1;
        score = gap_row;
      }
      if (max_score - score > cutoff) {
        if (array_start + 1 == ai) ++array_start;
        dp[ai].best = kInitScore;
        dp[ai].best_gap = kInitScore;
        gap_row = kInitScore;
      } else {
        last = ai;
        if (score > max_score) {
          max_score = score;
          max_p0 = p0;
          max_p1 = p1;
        }
        if (kTrace) {
          gap_col -= gap_ext;
          gap_row -= gap_ext;
          if (gap_col < score - gap_init) {
            dp[ai].best_gap = score - gap_init;
          } else {
            dp[ai].best_gap = gap_col;
            op += kOpExt0;
          }
          if (gap_row < score - gap_init) {
            gap_row = score - gap_init;
          } else {
            op += kOpExt1;
          }
          arena->Write(arr, row_zero + ai, op);
        } else {
          dp[ai].best_gap = std::max(score - gap_init, gap_col - gap_ext);
          gap_row = std::max(score - gap_init, gap_row - gap_ext);
        }
        dp[ai].best = score;
      }
    }
    if (array_start + 1 == array_end) break;
    if (last < array_end - 1) {
      array_end = last + 1;
    } else {
      while (gap_row >= (max_score - cutoff) && s0[p0] != delim) {
        dp[array_end].best = gap_row;
        dp[array_end].best_gap = gap_row - gap_init;
        if (kTrace) arena->Write(arr, row_zero + array_end, kOpGap0 + kOpExt0);
        gap_row -= gap_ext;
        ++array_end;
        p0 += inc;
      }
    }
    if (s0[p0] != delim) {
      dp[array_end].best = kInitScore;
      dp[array_end].best_gap = kInitScore;
      ++array_end;
    }
    if (kTrace) arena->st[arr].used += (uint32_t)(array_end - row_off);
  }
  *best_score = max_score;
  *best_p0 = max_p0;
  *best_p1 = max_p1;
  if (cells) *cells = ncells;

  if (kTrace) {
    // PickOptimalPath (src/gapped_extender.cpp:473-531)
    int q = max_p0, d = max_p1;
    uint8_t op = kOpSub, next_op = kOpSub;
    while (d * inc >= 0 && q * inc >= 0) {
      const int r = d * inc;
      next_op = arena->Read(row_arr[r], row_base[r] + (q * inc) - row_offset[r] + 1);
      switch (op) {
        case kOpGap0:
          op = next_op & 3;
          if (next_op & kOpExt0) op = kOpGap0;
          break;
        case kOpGap1:
          op = next_op & 3;
          if (next_op & kOpExt1) op = kOpGap1;
          break;
        default:
          op = next_op & 3;
          break;
      }
      switch (op) {
        case kOpGap0:
          eb->Add(kEditGap0, 1);
          d -= inc;
          break;
        case kOpGap1:
          eb->Add(kEditGap1, 1);
          q -= inc;
          break;
        default:
          eb->Add(kEditSub, 1);
          d -= inc;
          q -= inc;
          break;
      }
    }
    if (d * inc >= 0) {
      eb->Add(kEditGap0, d * inc + 1);
    } else if (q * inc >= 0) {
      eb->Add(kEditGap1, q * inc + 1);
    }
    if (inc > 0) eb->Reverse();
  }
  return true;
}

// ChainFilter (src/chain_filter.cpp:24-99)
struct Hit {
  uint32_t q;
  uint32_t d;
};
struct DbStartComp {
  bool operator()(const Hit &a, const Hit &b) const { return a.d < b.d; }
};

bool ChainConnected(const uint8_t *qs, const uint8_t *ds, uint8_t delim,
                    const int *sm, uint32_t nl, uint32_t chain_query_end,
                    const Hit &hit, int cutoff) {
  if (hit.q <= chain_query_end + 1) return true;
  uint32_t qo = hit.q, dof = hit.d;
  int score = 0, best = 0, c = -cutoff;
  do {
    --dof;
    --qo;
    if (ds[dof] == delim || qs[qo] == delim || qo <= chain_query_end) break;
    score += sm[ds[dof] * nl + qs[qo]];
    if (best < score) {
      best = score;
      c = best - cutoff;
    }
  } while (score > c);
  return qo <= chain_query_end;
}

void ChainFilterRun(const uint8_t *qs, uint32_t qlen, int cutoff,
                    const uint8_t *ds, uint8_t delim, const int *sm,
                    uint32_t nl, std::vector<Hit> &hits) {
  struct Chain {
    uint32_t query_end;
    uint32_t diagonal;
  };
  std::vector<Chain> chains(qlen);
  for (size_t i = 0; i < chains.size(); ++i) {
    chains[i].diagonal = 0;
    chains[i].query_end = 0;
  }
  std::sort(hits.begin(), hits.end(), DbStartComp());
  uint32_t n = 0;
  for (size_t i = 0; i < hits.size(); ++i) {
    const Hit h = hits[i];
    uint32_t diagonal = qlen - 1 + h.d - h.q;
    uint32_t off = diagonal % qlen;
    if (chains[off].diagonal != diagonal ||
        !ChainConnected(qs, ds, delim, sm, nl, chains[off].query_end, h,
                        cutoff)) {
      chains[off].query_end = h.q;
      chains[off].diagonal = qlen - 1 + h.d - h.q;
      hits[n] = h;
      ++n;
    } else {
      chains[off].query_end = h.q;
    }
  }
  hits.resize(n);
}

// ReducedAlphabetVariableHashFunction::CalculateHash
// (src/reduced_alphabet_variable_hash_function.cpp:56-86)
int HashAt(const uint8_t *s, const uint8_t *map, uint32_t map_size,
           const int *scores, uint32_t max_code, int threshold,
           uint32_t max_length, uint32_t shift, uint32_t *hash,
           uint32_t *hlen) {
  uint32_t h = 0;
  int score = 0;
  uint32_t len = 0;
  for (len = 0; len < max_length; ++len) {
    if (s[len] < map_size && map[s[len]] <= max_code) {
      if (len >= 6 && score >= threshold) break;  // kMinHashLength = 6
      h <<= shift;
      uint8_t c = map[s[len]];
      h += c;
      score += scores[c];
    } else {
      *hash = 0;
      *hlen = 0;
      return 1;
    }
  }
  *hash = h;
  *hlen = len;
  return 0;
}

// ---------------------------------------------------------------------------
// Statistics (src/statistics.cpp:28-49, 51-129)
// ---------------------------------------------------------------------------
struct Karlin {
  float lambda;
  float K;
  float H;
};

int Normalized2Nominal(float normalized, const Karlin &k) {
  float s = (normalized * static_cast<float>(std::log(2.0)) + std::log(k.K)) /
            k.lambda;
  return static_cast<int>(std::floor(s));
}
int NormalizedCutoff2NominalCutoff(float normalized, const Karlin &k) {
  float s = (normalized * static_cast<float>(std::log(2.0))) / k.lambda;
  return static_cast<int>(std::floor(s));
}
float Nominal2Normalized(int score, const Karlin &k) {
  return ((static_cast<float>(score) * k.lambda) - std::log(k.K)) /
         static_cast<float>(std::log(2.0));
}
double Nominal2EValue(int score, uint64_t space, const Karlin &k) {
  return space * k.K * std::exp(static_cast<double>(-1.0 * score * k.lambda));
}
uint64_t LengthAdjustment(uint64_t query_length, uint64_t db_length,
                          uint64_t db_n, const Karlin &p, double alpha,
                          double beta) {
  uint64_t adj = 0;
  const int kMaxIteration = 20;
  double m = static_cast<double>(query_length);
  double n = static_cast<double>(db_length);
  double N = static_cast<double>(db_n);
  double log_k = std::log(p.K);
  double adl = static_cast<double>(alpha / p.lambda);
  double e, e_min = 0, e_max;
  double ss;
  double a = N;
  double mb = m * N + n;
  double c = n * m - std::max(m, n) / p.K;
  if (c < 0) {
    return 0;
  } else {
    e_max = 2 * c / (mb + std::sqrt((mb * mb - 4 * a * c)));
  }
  bool converged = false;
  double e_next = 0.0;
  for (int i = 0; i < kMaxIteration; ++i) {
    double e_bar;
    e = e_next;
    ss = (m - e) * (n - N * e);
    e_bar = adl * (log_k + std::log(ss)) + beta;
    if (e_bar >= e) {
      e_min = e;
      if (e_bar - e_min <= 1.0) {
        converged = true;
        break;
      }
      if (e_min >= e_max) break;
    } else {
      e_max = e;
    }
    if (e_min <= e_bar && e_bar <= e_max) {
      e_next = e_bar;
    } else {
      e_next = (i == 0) ? e_max : (e_min + e_max) / 2;
    }
  }
  if (converged) {
    adj = static_cast<int>(e_min);
    e = std::ceil(e_min);
    if (e <= e_max) {
      ss = (m - e) * (n - N * e);
      if (adl * (log_k + std::log(ss)) + beta > e) adj = static_cast<int>(e);
    }
  } else {
    adj = static_cast<int>(e_min);
  }
  return adj;
}
uint64_t SearchSpace(uint64_t query_length, uint64_t db_length, uint64_t db_n,
                     const Karlin &p, double alpha, double beta) {
  uint64_t adj = LengthAdjustment(query_length, db_length, db_n, p, alpha, beta);
  uint64_t eff_db = db_length - (db_n * adj);
  return eff_db * (query_length - adj);
}

// ---------------------------------------------------------------------------
// The aligner
// ---------------------------------------------------------------------------
struct Coordinate {
  uint32_t q;
  uint32_t d;
};
struct Presearched {
  int chunk;
  uint32_t subject;
  int score;
  Coordinate hit, start, end;
  uint32_t hit_count;
};
struct Result {
  uint32_t chunk;
  uint32_t subject;
  int score;
  uint32_t length, mismatches, gap_openings, gap_extensions;
  Coordinate hit, start, end;
};

// One result of BuildResults (aligner_build_results_thread.h:97-150): both
// sides re-extended with traceback through the extender's arena (reverse
// side first), then AlignerCommon::BuildResult (src/aligner_common.h:276-343)
// over the joined edit script.
void BuildOneResult(const uint8_t *qs, uint32_t qlen, const uint8_t *ds, Coordinate hit,
                    uint8_t delim, const int *sm, uint32_t nl, int gap_open, int gap_ext,
                    int result_cutoff, GapArena *arena, Result *out) {
  EditBlocks eb, tmp;
  int sum = 0, score, p0, p1;
  Coordinate start, end;
  GappedOneSide<true>(qs + hit.q - 1, hit.q - 1, ds + hit.d - 1, delim, true, sm, nl, gap_open,
                      gap_ext, result_cutoff, &score, &p0, &p1, &tmp, NULL, arena);
  sum += score;
  start.q = hit.q - 1 + p0;
  start.d = hit.d - 1 + p1;
  tmp.Reverse();
  eb.Add(tmp);
  GappedOneSide<true>(qs + hit.q, qlen - hit.q, ds + hit.d, delim, false, sm, nl, gap_open,
                      gap_ext, result_cutoff, &score, &p0, &p1, &tmp, NULL, arena);
  sum += score;
  end.q = hit.q + p0;
  end.d = hit.d + p1;
  eb.Add(tmp);
  std::vector<int> edits = eb.ToVector();
  uint32_t qp = start.q, dp = start.d, m = 0, o = 0, e = 0;
  int prev = kEditSub;
  for (size_t x = 0; x < edits.size(); ++x) {
    switch (edits[x]) {
      case kEditSub:
        if (qs[qp] != ds[dp]) ++m;
        ++qp;
        ++dp;
        break;
      case kEditGap0:
        if (prev != kEditGap0) ++o; else ++e;
        ++dp;
        break;
      case kEditGap1:
        if (prev != kEditGap1) ++o; else ++e;
        ++qp;
        break;
    }
    prev = edits[x];
  }
  out->score = sum;
  out->length = (uint32_t)edits.size();
  out->mismatches = m;
  out->gap_openings = o;
  out->gap_extensions = e;
  out->hit = hit;
  out->start = start;
  out->end = end;
}

// AlignerCommon::AlignmentComp (src/aligner_common.cpp:12-56)
int AlignmentComp(int as, const Coordinate &a0, const Coordinate &a1, int bs,
                  const Coordinate &b0, const Coordinate &b1) {
  if (as > bs) return 1;
  if (as < bs) return -1;
  uint32_t aq = a1.q - a0.q, bq = b1.q - b0.q;
  if (aq > bq) return 1;
  if (aq < bq) return -1;
  uint32_t ad = a1.d - a0.d, bd = b1.d - b0.d;
  if (ad > bd) return 1;
  if (ad < bd) return -1;
  if (a0.d < b0.d) return 1;
  return -1;
}
bool IsContained(const Coordinate &a0, const Coordinate &a1,
                 const Coordinate &b0, const Coordinate &b1) {
  return (a0.q == b0.q && a0.d == b0.d) || (a1.q == b1.q && a1.d == b1.d) ||
         (a0.q <= b0.q && b1.q <= a1.q && a0.d <= b0.d && b1.d <= a1.d);
}
struct PresearchedGreater {
  bool operator()(const Presearched &a, const Presearched &b) const {
    return AlignmentComp(a.score, a.start, a.end, b.score, b.start, b.end) > 0;
  }
};
struct ResultGreater {
  bool operator()(const Result &a, const Result &b) const {
    return AlignmentComp(a.score, a.start, a.end, b.score, b.start, b.end) > 0;
  }
};

struct ChunkView {
  const gzo_db_chunk *c;
  int id;
  // DatabaseChunk::GetId (src/database_chunk.h:214-240): the subject that
  // contains `position`.
  uint32_t GetId(uint32_t position) const {
    const uint32_t *o = c->offsets;
    uint32_t n = c->number_sequences;
    const uint32_t *it = std::upper_bound(o, o + n, position);
    return (uint32_t)(it - o) - 1;
  }
  bool Lookup(uint32_t hash, uint32_t *idx) const {
    const uint32_t *h = c->hashes;
    const uint32_t *it = std::lower_bound(h, h + c->number_hashes, hash);
    if (it == h + c->number_hashes || *it != hash) return false;
    *idx = (uint32_t)(it - h);
    return true;
  }
};

struct AlignmentPosition {
  uint32_t result_id;
  uint32_t position;
};
struct AlignmentPositionLess {
  bool operator()(const AlignmentPosition &a,
                  const AlignmentPosition &b) const {
    return a.position < b.position;
  }
};

// AlignerCommon::AddResults (src/aligner_common.h:147-274)
void AddResults(const ChunkView &db, const gzo_parameters &p,
                std::vector<std::vector<AlignmentPosition> > &starts,
                std::vector<Presearched> &added,
                std::vector<Presearched> &results) {
  std::vector<Presearched> nr;
  if (starts.size() < db.c->number_sequences) starts.resize(db.c->number_sequences);
  std::stable_sort(added.begin(), added.end(), PresearchedGreater());
  size_t ri = 0, ai = 0;
  while (nr.size() < p.max_results &&
         (ai != added.size() || ri != results.size())) {
    Presearched a;
    if (ai == added.size()) {
      a = results[ri++];
    } else if (ri == results.size() ||
               AlignmentComp(results[ri].score, results[ri].start,
                             results[ri].end, added[ai].score, added[ai].start,
                             added[ai].end) < 0) {
      a.chunk = db.id;
      a.subject = db.GetId(added[ai].start.d);
      a.score = added[ai].score;
      a.start = added[ai].start;
      a.end = added[ai].end;
      a.hit = added[ai].hit;
      a.hit_count = added[ai].hit_count;
      ++ai;
    } else {
      a = results[ri++];
    }
    if (a.chunk == db.id) {
      std::vector<AlignmentPosition> &sp = starts[a.subject];
      if (sp.size() < p.max_results_per_subject) {
        AlignmentPosition key;
        key.result_id = 0;
        key.position = a.start.d;
        std::vector<AlignmentPosition>::iterator it =
            std::lower_bound(sp.begin(), sp.end(), key, AlignmentPositionLess());
        std::vector<AlignmentPosition>::iterator ins = it;
        key.position = a.end.d;
        std::vector<AlignmentPosition>::iterator ie =
            std::lower_bound(it, sp.end(), key, AlignmentPositionLess());
        bool ok = true;
        for (; it != ie; ++it) {
          Presearched *r = &nr[it->result_id];
          if (IsContained(r->start, r->end, a.start, a.end)) {
            ok = false;
            break;
          }
        }
        if (ok) {
          AlignmentPosition ap;
          ap.position = a.start.q;
          ap.result_id = (uint32_t)nr.size();
          sp.insert(ins, ap);
          nr.push_back(a);
        }
      }
    } else {
      nr.push_back(a);
    }
  }
  results = nr;
  for (size_t i = 0; i < nr.size(); ++i) starts[nr[i].subject].clear();
}

// per thread, so that searches on several threads at once keep their own counts
thread_local gzo_stats g_stats;
// 0: presearch end = hit + forward best offset; 1: end zero-initialised and
// accumulated over one query's hits of one chunk
thread_local int g_end_reading = 0;

struct ListEntry {
  bool check;
  uint32_t query_id;
  uint32_t position;
};
struct HashList {
  uint32_t hash;
  std::vector<ListEntry> data;
};

// SeedSearcherQueryParameters::Build with one thread
// (src/seed_searcher_query_parameters.cpp:17-219): hash lists in order of
// first appearance, queries popped in reverse within blocks of 64.
void BuildHashLists(const gzo_queries &Q, const gzo_database &D,
                    uint8_t delim, std::vector<HashList> &lists) {
  std::unordered_map<uint32_t, uint32_t> ids;
  const uint32_t L = D.subsequence_length;
  for (uint32_t block = 0; block < Q.number_queries; block += 64) {
    uint32_t end = std::min(Q.number_queries, block + 64);
    for (uint32_t qi = end; qi-- > block;) {
      const uint8_t *s = Q.sequences + Q.offsets[qi];
      uint32_t len = (uint32_t)(Q.offsets[qi + 1] - Q.offsets[qi]);
      uint32_t off = 0;
      while (off < len) {
        uint32_t e = off;
        while (e < len && s[e] != delim) ++e;
        if (e - off >= L) {
          uint32_t cs = off + L / 2, ce = e - L / 2;
          for (uint32_t i = off; i < e; ++i) {
            uint32_t h, hl;
            if (!HashAt(s + i, D.hash_code_map, 25, D.hash_code_scores,
                        D.hash_max_code, D.hash_score_threshold,
                        D.hash_max_length, D.hash_shift_size, &h, &hl)) {
              ListEntry le;
              le.query_id = qi;
              le.position = i;  // CalculateSeedPosition returns the start
              le.check = (cs <= i && i < ce);
              std::unordered_map<uint32_t, uint32_t>::iterator f = ids.find(h);
              uint32_t id;
              if (f == ids.end()) {
                id = (uint32_t)lists.size();
                ids[h] = id;
                lists.push_back(HashList());
                lists.back().hash = h;
              } else {
                id = f->second;
              }
              lists[id].data.push_back(le);
            }
          }
        }
        off = e + 1;
      }
    }
  }
}

// The seed search of one database chunk for every query: hits appended per
// query in the one-thread enumeration order (AlignerPresearchThread::Run,
// src/aligner_presearch_thread.cpp:140-152, with SeedSearcher::Search,
// src/seed_searcher.cpp:103-283).
void SearchChunk(const gzo_queries &Q, const gzo_database &D, const ChunkView &db,
                 const std::vector<HashList> &lists, char distance_threshold,
                 int max_mismatch, uint8_t delim, const int *sm, uint32_t nl,
                 const std::vector<int> &ucut, const std::vector<int> &trig,
                 std::vector<std::vector<Hit> > &hits) {
  const uint32_t L = D.subsequence_length;
  const uint8_t *ds = db.c->sequence;
  // AlignerPresearchThread::Run seed-search loop (:140-152): blocks of
  // 256 hash lists, popped from the back.
  const uint32_t nlists = (uint32_t)lists.size();
  for (uint32_t block = 0; block < nlists; block += 256) {
    uint32_t bend = std::min(nlists, block + 256);
    for (uint32_t li = bend; li-- > block;) {
      const HashList &hl = lists[li];
      uint32_t hidx;
      if (!db.Lookup(hl.hash, &hidx)) continue;
      // SearchFromClusteringSubsequences (src/seed_searcher.cpp:103-223)
      if (db.c->clustering) {
        for (uint32_t r = db.c->rep_offsets[hidx];
             r < db.c->rep_offsets[hidx + 1]; ++r) {
          const uint32_t rpos = db.c->rep_positions[r];
          std::vector<uint32_t> hitting;
          for (size_t k = 0; k < hl.data.size(); ++k) {
            const ListEntry &e = hl.data[k];
            const uint8_t *qs = Q.sequences + Q.offsets[e.query_id];
            int dist = 0;
            if (e.check) {
              ++g_stats.distance_calculations;
              dist = Distance(qs, e.position, ds, rpos, L, D.reduced_code_map);
              int lower = std::max(dist - max_mismatch, 0);
              if (distance_threshold >= lower) hitting.push_back((uint32_t)k);
            } else {
              hitting.push_back((uint32_t)k);
            }
            if (distance_threshold >= dist) {
              ++g_stats.ungapped_extensions;
              int s = UngappedWithTrigger(qs, e.position, ds, rpos, delim,
                                          sm, nl, ucut[e.query_id],
                                          trig[e.query_id]);
              if (s > trig[e.query_id]) {
                Hit h;
                h.q = e.position;
                h.d = rpos;
                hits[e.query_id].push_back(h);
              }
            }
          }
          if (!hitting.empty()) {
            for (uint32_t m = db.c->rep_member_offsets[r];
                 m < db.c->rep_member_offsets[r + 1]; ++m) {
              const uint32_t mpos = db.c->member_positions[m];
              for (size_t k = 0; k < hitting.size(); ++k) {
                const ListEntry &e = hl.data[hitting[k]];
                const uint8_t *qs = Q.sequences + Q.offsets[e.query_id];
                ++g_stats.ungapped_extensions;
                int s = UngappedWithTrigger(qs, e.position, ds, mpos, delim,
                                            sm, nl, ucut[e.query_id],
                                            trig[e.query_id]);
                if (s > trig[e.query_id]) {
                  Hit h;
                  h.q = e.position;
                  h.d = mpos;
                  hits[e.query_id].push_back(h);
                }
              }
            }
          }
        }
      }
      // SearchFromNonClusteringSubsequences (src/seed_searcher.cpp:225-283)
      for (uint32_t v = db.c->value_offsets[hidx];
           v < db.c->value_offsets[hidx + 1]; ++v) {
        const uint32_t vpos = db.c->value_positions[v];
        for (size_t k = 0; k < hl.data.size(); ++k) {
          const ListEntry &e = hl.data[k];
          const uint8_t *qs = Q.sequences + Q.offsets[e.query_id];
          ++g_stats.ungapped_extensions;
          int s = UngappedWithTrigger(qs, e.position, ds, vpos, delim, sm,
                                      nl, ucut[e.query_id], trig[e.query_id]);
          if (s > trig[e.query_id]) {
            Hit h;
            h.q = e.position;
            h.d = vpos;
            hits[e.query_id].push_back(h);
          }
        }
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int gzo_ungapped_one_side(const uint8_t *seq0, const uint8_t *seq1,
                          uint8_t delimiter, int reversed, const int *matrix,
                          uint32_t number_letters, int cutoff) {
  return UngappedScoreOnly(seq0, seq1, delimiter, reversed != 0, matrix,
                           number_letters, cutoff);
}

int gzo_ungapped_one_side_positions(const uint8_t *seq0, const uint8_t *seq1,
                                    uint8_t delimiter, int reversed,
                                    const int *matrix, uint32_t number_letters,
                                    int cutoff, int *p0, int *p1) {
  return UngappedWithPositions(seq0, seq1, delimiter, reversed != 0, matrix,
                               number_letters, cutoff, p0, p1);
}

int gzo_ungapped_with_trigger(const uint8_t *query, uint32_t qpos,
                              const uint8_t *database, uint32_t dpos,
                              uint8_t delimiter, const int *matrix,
                              uint32_t number_letters, int cutoff,
                              int trigger) {
  return UngappedWithTrigger(query, qpos, database, dpos, delimiter, matrix,
                             number_letters, cutoff, trigger);
}

int gzo_distance(const uint8_t *query, uint32_t qc, const uint8_t *database,
                 uint32_t dc, uint32_t L, const uint8_t *map) {
  return Distance(query, qc, database, dc, L, map);
}

int gzo_gapped_one_side(const uint8_t *seq0, uint32_t seq0_length,
                        const uint8_t *seq1, uint8_t delimiter, int reversed,
                        const int *matrix, uint32_t number_letters,
                        int gap_open, int gap_extension, int cutoff,
                        int *best_score, int *p0, int *p1, uint8_t *edits,
                        uint32_t *number_edits, uint64_t *dp_cells) {
  if (edits == NULL) {
    GappedOneSide<false>(seq0, seq0_length, seq1, delimiter, reversed != 0,
                         matrix, number_letters, gap_open, gap_extension,
                         cutoff, best_score, p0, p1, NULL, dp_cells, NULL);
    if (number_edits) *number_edits = 0;
    return 0;
  }
  // a fresh extender object, as a single reference call would use
  EditBlocks eb;
  GapArena arena;
  GappedOneSide<true>(seq0, seq0_length, seq1, delimiter, reversed != 0,
                      matrix, number_letters, gap_open, gap_extension, cutoff,
                      best_score, p0, p1, &eb, dp_cells, &arena);
  std::vector<int> v = eb.ToVector();
  for (size_t i = 0; i < v.size(); ++i) edits[i] = (uint8_t)v[i];
  if (number_edits) *number_edits = (uint32_t)v.size();
  return 0;
}

int gzo_build_chunk(const uint8_t *queries, const uint64_t *query_offsets,
                    uint32_t number_queries, const uint8_t *chunk, const uint32_t *hits,
                    uint32_t number_hits, uint8_t delimiter, const int *matrix, uint32_t nl,
                    int gap_open, int gap_extension, int result_cutoff, int32_t *out) {
  (void)number_queries;
  // BuildResults of one chunk (src/aligner.cpp:264-348,
  // aligner_build_results_thread.h:64-190): the results in query order,
  // handed out in blocks of 32 and popped from the back of each block, all
  // through one extender object
  GapArena arena;
  for (uint32_t block = 0; block < number_hits; block += 32) {
    const uint32_t bend = std::min(number_hits, block + 32);
    for (uint32_t k = bend; k-- > block;) {
      const uint32_t qi = hits[3 * k];
      const uint8_t *qs = queries + query_offsets[qi];
      const uint32_t qlen = (uint32_t)(query_offsets[qi + 1] - query_offsets[qi]);
      Coordinate hit;
      hit.q = hits[3 * k + 1];
      hit.d = hits[3 * k + 2];
      Result r;
      BuildOneResult(qs, qlen, chunk, hit, delimiter, matrix, nl, gap_open, gap_extension,
                     result_cutoff, &arena, &r);
      int32_t *o = out + 9 * (size_t)k;
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

uint32_t gzo_chain_filter(const uint8_t *query, uint32_t qlen, int cutoff,
                          const uint8_t *database, uint8_t delimiter,
                          const int *matrix, uint32_t nl, uint32_t *hits,
                          uint32_t nh) {
  std::vector<Hit> v(nh);
  for (uint32_t i = 0; i < nh; ++i) {
    v[i].q = hits[2 * i];
    v[i].d = hits[2 * i + 1];
  }
  ChainFilterRun(query, qlen, cutoff, database, delimiter, matrix, nl, v);
  for (size_t i = 0; i < v.size(); ++i) {
    hits[2 * i] = v[i].q;
    hits[2 * i + 1] = v[i].d;
  }
  return (uint32_t)v.size();
}

int gzo_hash(const uint8_t *sequence, const uint8_t *code_map,
             uint32_t code_map_size, const int *code_scores, uint32_t max_code,
             int score_threshold, uint32_t max_length, uint32_t shift_size,
             uint32_t *hash, uint32_t *hash_length) {
  return HashAt(sequence, code_map, code_map_size, code_scores, max_code,
                score_threshold, max_length, shift_size, hash, hash_length);
}

void gzo_last_stats(gzo_stats *s) { *s = g_stats; }

void gzo_set_end_reading(int reading) { g_end_reading = reading; }

int gzo_query_lists(const gzo_queries *Qp, const gzo_database *Dp, uint64_t *number_lists,
                    uint32_t **hash, uint32_t **entry_offsets, uint64_t *number_entries,
                    uint32_t **entry_queries, uint32_t **entry_positions,
                    uint8_t **entry_checks) {
  std::vector<HashList> lists;
  BuildHashLists(*Qp, *Dp, 25, lists);
  // the presearch visits blocks of 256 lists, each from its back
  // (AlignerPresearchThread::Run, src/aligner_presearch_thread.cpp:140-152)
  const uint32_t nl = (uint32_t)lists.size();
  std::vector<uint32_t> visit;
  for (uint32_t block = 0; block < nl; block += 256) {
    const uint32_t bend = std::min(nl, block + 256);
    for (uint32_t li = bend; li-- > block;) visit.push_back(li);
  }
  uint64_t ne = 0;
  for (const HashList &l : lists) ne += l.data.size();
  *number_lists = nl;
  *number_entries = ne;
  *hash = (uint32_t *)malloc(sizeof(uint32_t) * (nl + 1));
  *entry_offsets = (uint32_t *)malloc(sizeof(uint32_t) * (nl + 1));
  *entry_queries = (uint32_t *)malloc(sizeof(uint32_t) * (ne + 1));
  *entry_positions = (uint32_t *)malloc(sizeof(uint32_t) * (ne + 1));
  *entry_checks = (uint8_t *)malloc(ne + 1);
  uint64_t o = 0;
  for (uint32_t r = 0; r < nl; ++r) {
    const HashList &l = lists[visit[r]];
    (*hash)[r] = l.hash;
    (*entry_offsets)[r] = (uint32_t)o;
    for (const ListEntry &e : l.data) {
      (*entry_queries)[o] = e.query_id;
      (*entry_positions)[o] = e.position;
      (*entry_checks)[o] = e.check ? 1 : 0;
      ++o;
    }
  }
  (*entry_offsets)[nl] = (uint32_t)o;
  return 0;
}

int gzo_add_results(const uint32_t *subject_offsets, uint32_t number_subjects, int chunk_id,
                    uint32_t max_results, uint32_t max_results_per_subject,
                    gzo_presearched *added, uint32_t number_added, gzo_presearched *results,
                    uint32_t *number_results) {
  gzo_db_chunk c;
  memset(&c, 0, sizeof(c));
  c.offsets = subject_offsets;
  c.number_sequences = number_subjects;
  ChunkView db;
  db.c = &c;
  db.id = chunk_id;
  gzo_parameters p;
  memset(&p, 0, sizeof(p));
  p.max_results = max_results;
  p.max_results_per_subject = max_results_per_subject;
  auto to = [](const gzo_presearched &x) {
    Presearched r;
    r.chunk = x.chunk;
    r.subject = x.subject;
    r.score = x.score;
    r.hit.q = x.hit_q;
    r.hit.d = x.hit_d;
    r.start.q = x.start_q;
    r.start.d = x.start_d;
    r.end.q = x.end_q;
    r.end.d = x.end_d;
    r.hit_count = x.hit_count;
    return r;
  };
  auto from = [](const Presearched &r) {
    gzo_presearched x;
    x.chunk = r.chunk;
    x.subject = r.subject;
    x.score = r.score;
    x.hit_q = r.hit.q;
    x.hit_d = r.hit.d;
    x.start_q = r.start.q;
    x.start_d = r.start.d;
    x.end_q = r.end.q;
    x.end_d = r.end.d;
    x.hit_count = r.hit_count;
    return x;
  };
  std::vector<Presearched> a, r;
  for (uint32_t i = 0; i < number_added; ++i) a.push_back(to(added[i]));
  for (uint32_t i = 0; i < *number_results; ++i) r.push_back(to(results[i]));
  std::vector<std::vector<AlignmentPosition> > starts;
  AddResults(db, p, starts, a, r);
  for (uint32_t i = 0; i < number_added; ++i) added[i] = from(a[i]);
  for (size_t i = 0; i < r.size(); ++i) results[i] = from(r[i]);
  *number_results = (uint32_t)r.size();
  return 0;
}

int gzo_seed_hits(const gzo_queries *Qp, const gzo_database *Dp, const gzo_parameters *Pp,
                  uint32_t chunk, uint32_t **query_offsets, uint32_t **hits) {
  const gzo_queries &Q = *Qp;
  const gzo_database &D = *Dp;
  const gzo_parameters &P = *Pp;
  if (chunk >= D.number_chunks) return 1;
  const uint32_t nq = Q.number_queries;
  std::vector<int> ucut(nq), trig(nq);
  for (uint32_t i = 0; i < nq; ++i) {
    Karlin k;
    k.lambda = Q.lambda[i];
    k.K = Q.K[i];
    k.H = 0;
    ucut[i] = NormalizedCutoff2NominalCutoff(P.normalized_ungapped_cutoff, k);
    trig[i] = Normalized2Nominal(P.normalized_gapped_trigger, k);
  }
  std::vector<HashList> lists;
  BuildHashLists(Q, D, P.delimiter, lists);
  const uint32_t L = D.subsequence_length;
  const float rate = 0.8f;
  const char distance_threshold = L - rate * L;
  ChunkView db;
  db.c = &D.chunks[chunk];
  db.id = (int)chunk;
  std::vector<std::vector<Hit> > hv(nq);
  SearchChunk(Q, D, db, lists, distance_threshold, D.max_mismatches, P.delimiter, P.matrix,
              P.number_letters, ucut, trig, hv);
  size_t total = 0;
  for (const auto &v : hv) total += v.size();
  *query_offsets = (uint32_t *)malloc(sizeof(uint32_t) * (nq + 1));
  *hits = (uint32_t *)malloc(sizeof(uint32_t) * (2 * total + 1));
  size_t o = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    (*query_offsets)[i] = (uint32_t)o;
    for (const Hit &h : hv[i]) {
      (*hits)[2 * o] = h.q;
      (*hits)[2 * o + 1] = h.d;
      ++o;
    }
  }
  (*query_offsets)[nq] = (uint32_t)o;
  return 0;
}

void gzo_free(void *p) { free(p); }

int gzo_align(const gzo_queries *Qp, const gzo_database *Dp,
              const gzo_parameters *Pp, char **output, size_t *output_length) {
  const gzo_queries &Q = *Qp;
  const gzo_database &D = *Dp;
  const gzo_parameters &P = *Pp;
  const uint8_t delim = P.delimiter;
  const int *sm = P.matrix;
  const uint32_t nl = P.number_letters;
  memset(&g_stats, 0, sizeof(g_stats));

  // gapped Karlin for BLOSUM62 11/1 (src/statistics.cpp:236-249)
  Karlin gk;
  gk.lambda = 0.267f;
  gk.K = 0.041f;
  gk.H = 0.14f;
  const int presearch_cutoff =
      Normalized2Nominal(P.normalized_presearch_cutoff, gk);
  const int result_cutoff = Normalized2Nominal(P.normalized_result_cutoff, gk);

  // Aligner::SetQueriesData (src/aligner.cpp:247-262)
  const uint32_t nq = Q.number_queries;
  std::vector<int> ucut(nq), trig(nq);
  for (uint32_t i = 0; i < nq; ++i) {
    Karlin k;
    k.lambda = Q.lambda[i];
    k.K = Q.K[i];
    k.H = 0;
    ucut[i] = NormalizedCutoff2NominalCutoff(P.normalized_ungapped_cutoff, k);
    trig[i] = Normalized2Nominal(P.normalized_gapped_trigger, k);
  }

  std::vector<HashList> lists;
  BuildHashLists(Q, D, delim, lists);
  g_stats.hash_lists = lists.size();

  // SeedSearcher::Reset (src/seed_searcher.cpp:47-78)
  const uint32_t L = D.subsequence_length;
  const float rate = 0.8f;
  const char distance_threshold = L - rate * L;
  const int max_mismatch = D.max_mismatches;

  std::vector<std::vector<Presearched> > presearch(nq);
  std::vector<std::vector<Hit> > hits(nq);
  std::vector<std::vector<AlignmentPosition> > starts;

  for (uint32_t ci = 0; ci < D.number_chunks; ++ci) {
    ChunkView db;
    db.c = &D.chunks[ci];
    db.id = (int)ci;
    const uint8_t *ds = db.c->sequence;

    SearchChunk(Q, D, db, lists, distance_threshold, max_mismatch, delim, sm, nl, ucut, trig,
                hits);

    // chain filter + gapped extension + AddResults per query
    // (src/aligner_presearch_thread.cpp:157-206, Extend :315-383)
    for (uint32_t qi = 0; qi < nq; ++qi) {
      std::vector<Hit> &hv = hits[qi];
      g_stats.hits += hv.size();
      const uint8_t *qs = Q.sequences + Q.offsets[qi];
      const uint32_t qlen = (uint32_t)(Q.offsets[qi + 1] - Q.offsets[qi]);
      ChainFilterRun(qs, qlen, ucut[qi], ds, delim, sm, nl, hv);
      g_stats.hits_after_chain_filter += hv.size();
      std::vector<Presearched> tmp;
      Coordinate acc = {0, 0};
      for (size_t i = 0; i < hv.size(); ++i) {
        Presearched r;
        r.chunk = -1;
        r.subject = 0;
        r.score = 0;
        r.hit.q = hv[i].q;
        r.hit.d = hv[i].d;
        int score, p0, p1;
        GappedOneSide<false>(qs + r.hit.q - 1, r.hit.q - 1, ds + r.hit.d - 1,
                             delim, true, sm, nl, P.gap_open, P.gap_extension,
                             presearch_cutoff, &score, &p0, &p1, NULL, NULL, NULL);
        r.score += score;
        r.start.q = r.hit.q - 1 + p0;
        r.start.d = r.hit.d - 1 + p1;
        GappedOneSide<false>(qs + r.hit.q, qlen - r.hit.q, ds + r.hit.d, delim,
                             false, sm, nl, P.gap_open, P.gap_extension,
                             presearch_cutoff, &score, &p0, &p1, NULL, NULL, NULL);
        r.score += score;
        if (g_end_reading == 0) {
          r.end.q = r.hit.q + p0;
          r.end.d = r.hit.d + p1;
        } else {
          // AlignerPresearchThread::Extend (:322, 363-364): one tmp_result
          // per call, end += offset + 1 for each hit
          acc.q += p0 + 1;
          acc.d += p1 + 1;
          r.end = acc;
        }
        r.hit_count = 1;
        tmp.push_back(r);
        g_stats.gapped_extensions += 2;
      }
      AddResults(db, P, starts, tmp, presearch[qi]);
      hv.clear();
    }
  }

  // BuildResults (src/aligner.cpp:264-348, aligner_build_results_thread.h
  // :64-190): per chunk, (query, result) pairs in query order, handed out
  // in blocks of 32 and popped from the back of each block, all through one
  // GappedExtender object per chunk.
  std::vector<std::vector<Result> > results(nq);
  for (uint32_t qi = 0; qi < nq; ++qi) results[qi].resize(presearch[qi].size());
  for (uint32_t ci = 0; ci < D.number_chunks; ++ci) {
    std::vector<std::pair<uint32_t, uint32_t> > ids;
    for (uint32_t qi = 0; qi < nq; ++qi)
      for (uint32_t ri = 0; ri < presearch[qi].size(); ++ri)
        if (presearch[qi][ri].chunk == (int)ci) ids.push_back(std::make_pair(qi, ri));
    GapArena arena;
    const uint8_t *ds = D.chunks[ci].sequence;
    for (size_t block = 0; block < ids.size(); block += 32) {
      size_t bend = std::min(ids.size(), block + 32);
      for (size_t k = bend; k-- > block;) {
        const uint32_t qi = ids[k].first, ri = ids[k].second;
        const uint8_t *qs = Q.sequences + Q.offsets[qi];
        const uint32_t qlen = (uint32_t)(Q.offsets[qi + 1] - Q.offsets[qi]);
        const Presearched &pr = presearch[qi][ri];
        Result &r = results[qi][ri];
        BuildOneResult(qs, qlen, ds, pr.hit, delim, sm, nl, P.gap_open, P.gap_extension,
                       result_cutoff, &arena, &r);
        r.chunk = pr.chunk;
        r.subject = pr.subject;
      }
    }
    g_stats.stale_traceback_reads += arena.stale_reads;
  }
  for (uint32_t qi = 0; qi < nq; ++qi)
    std::sort(results[qi].begin(), results[qi].end(), ResultGreater());

  // AlignerCommon::WriteOutput (src/aligner_common.h:345-397)
  std::ostringstream os;
  const double alpha = 1.9, beta = -30;
  for (uint32_t qi = 0; qi < nq; ++qi) {
    const uint32_t qlen = (uint32_t)(Q.offsets[qi + 1] - Q.offsets[qi]);
    uint32_t real_length = Q.translated ? Q.dna_lengths[qi] / 3 : qlen - 2;
    uint64_t space = SearchSpace(real_length, D.total_length,
                                 D.total_sequences, gk, alpha, beta);
    for (size_t ri = 0; ri < results[qi].size(); ++ri) {
      const Result &r = results[qi][ri];
      const gzo_db_chunk &c = D.chunks[r.chunk];
      uint32_t qs, qe;
      if (Q.translated) {
        // TranslatedDnaQuery::GetRealStart/End (translated_dna_query.h:40-60)
        const uint32_t dl = Q.dna_lengths[qi];
        const uint32_t base = dl / 3 + 1;
        uint32_t f = (r.start.q - 1) / base;
        qs = (f < 3) ? ((r.start.q % base) - 1) * 3 + f + 1
                     : dl - (((r.start.q % base) - 1) * 3 + f - 2) + 1;
        f = (r.end.q - 1) / base;
        uint32_t rs = (f < 3) ? ((r.end.q % base) - 1) * 3 + f + 1
                              : dl - (((r.end.q % base) - 1) * 3 + f - 2) + 1;
        qe = (f < 3) ? rs + 2 : rs - 2;
      } else {
        qs = r.start.q;
        qe = r.end.q;
      }
      const uint32_t base_off = c.offsets[r.subject] - 1;
      os << Q.names[qi] << "\t";
      os << c.names[r.subject] << "\t";
      os << (1.0 - (static_cast<float>(r.mismatches + r.gap_openings +
                                       r.gap_extensions) /
                    static_cast<float>(r.length))) *
                100
         << "\t";
      os << r.length << "\t";
      os << r.mismatches << "\t";
      os << r.gap_openings << "\t";
      os << qs << "\t";
      os << qe << "\t";
      os << r.start.d - base_off << "\t";
      os << r.end.d - base_off << "\t";
      os << Nominal2EValue(r.score, space, gk) << "\t";
      os << Nominal2Normalized(r.score, gk);
      os << '\n';
    }
  }
  std::string s = os.str();
  char *buf = (char *)malloc(s.size() + 1);
  memcpy(buf, s.data(), s.size());
  buf[s.size()] = 0;
  *output = buf;
  *output_length = s.size();
  return 0;
}

}  // extern "C"
