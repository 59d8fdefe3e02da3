// database.cpp -- see database.h.
//
// The index is the one SeedSearcherDatabaseParameters::Build produces
// (src/seed_searcher_database_parameters.cpp:41-319):
//  1. every position is hashed with the variable-length reduced-alphabet
//     hash; a position whose hash is shorter than its predecessor's minus one
//     replaces that predecessor (:93-113);
//  2. per hash, positions are clustered greedily in position order: a
//     position not yet taken becomes a representative and takes every later
//     free position whose reduced window of L residues around the centre has
//     at most max_mismatches mismatches (:116-205); windows that touch a
//     delimiter or the chunk start never cluster;
//  3. representatives with members go to the cluster index, members to their
//     representative, everything else back to the plain index (:207-285).
// Within a hash bucket the last L - L/2 window residues are part of the
// (>= 6 residue) hash itself and therefore equal, so the distance reduces to
// the first L/2 residues; the greedy result is then "the earliest earlier
// representative within Hamming distance max_mismatches", found with L/2
// wildcard keys per position.  Other parameter sets use the direct greedy.
#include "database.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "alphabet.h"
#include "gz_common.h"

namespace gz {

uint32_t DbChunk::SubjectOf(uint32_t position) const {
  const uint32_t *o = offsets.data();
  const uint32_t n = NumberSequences();
  return (uint32_t)(std::upper_bound(o, o + n, position) - o) - 1;
}

namespace {

int ThreadCount(int requested) {
  int t = requested > 0 ? requested : (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(t, 64));
}

template <class F>
void ParallelFor(size_t n, int threads, size_t grain, F f) {
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

struct KV {
  uint32_t hash;
  uint32_t pos;
};

// Stable LSD radix sort of (hash, pos) by hash: two 16-bit passes.
void RadixSortByHash(std::vector<KV> &a, int threads) {
  const size_t n = a.size();
  if (n < 2) return;
  std::vector<KV> tmp(n);
  const int T = std::max(1, std::min<int>(threads, (int)(n / 65536) + 1));
  const size_t per = (n + T - 1) / T;
  std::vector<std::vector<size_t> > hist(T, std::vector<size_t>(65536));
  for (int pass = 0; pass < 2; ++pass) {
    const int shift = pass * 16;
    std::vector<KV> &src = pass == 0 ? a : tmp;
    std::vector<KV> &dst = pass == 0 ? tmp : a;
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
      pool.emplace_back([&, t]() {
        std::fill(hist[t].begin(), hist[t].end(), 0);
        const size_t b = t * per, e = std::min(n, b + per);
        for (size_t i = b; i < e; ++i) ++hist[t][(src[i].hash >> shift) & 0xffff];
      });
    }
    for (auto &p : pool) p.join();
    pool.clear();
    size_t sum = 0;
    for (int d = 0; d < 65536; ++d)
      for (int t = 0; t < T; ++t) {
        const size_t c = hist[t][d];
        hist[t][d] = sum;
        sum += c;
      }
    for (int t = 0; t < T; ++t) {
      pool.emplace_back([&, t]() {
        const size_t b = t * per, e = std::min(n, b + per);
        std::vector<size_t> &h = hist[t];
        for (size_t i = b; i < e; ++i) dst[h[(src[i].hash >> shift) & 0xffff]++] = src[i];
      });
    }
    for (auto &p : pool) p.join();
  }
}

// Open-addressing map masked key -> earliest representative slot.
struct KeyTable {
  std::vector<uint64_t> slots;  // (key << 32) | value, empty = ~0
  uint32_t mask = 0;
  void Reset(size_t entries) {
    size_t cap = 16;
    while (cap < entries * 2) cap <<= 1;
    if (slots.size() < cap) slots.resize(cap);
    mask = (uint32_t)cap - 1;
    std::fill(slots.begin(), slots.begin() + cap, ~0ull);
  }
  static uint32_t Mix(uint32_t k) {
    k ^= k >> 16;
    k *= 0x7feb352dU;
    k ^= k >> 15;
    return k;
  }
  bool Find(uint32_t key, uint32_t *value) const {
    for (uint32_t i = Mix(key) & mask;; i = (i + 1) & mask) {
      const uint64_t s = slots[i];
      if (s == ~0ull) return false;
      if ((uint32_t)(s >> 32) == key) {
        *value = (uint32_t)s;
        return true;
      }
    }
  }
  void InsertIfAbsent(uint32_t key, uint32_t value) {
    for (uint32_t i = Mix(key) & mask;; i = (i + 1) & mask) {
      const uint64_t s = slots[i];
      if (s == ~0ull) {
        slots[i] = ((uint64_t)key << 32) | value;
        return;
      }
      if ((uint32_t)(s >> 32) == key) return;
    }
  }
};

struct BucketOut {
  std::vector<uint32_t> hashes, rep_count, rep_pos, mem_count, mem_pos, val_count,
      val_pos;
};

}  // namespace

void BuildChunkIndex(DbChunk &chunk, const Database &common,
                     const DatabaseOptions &opt) {
  const int threads = ThreadCount(opt.threads);
  const uint8_t *seq = chunk.seq.data();
  const uint32_t length = (uint32_t)chunk.seq.size();
  const HashFunction &hf = common.hash;
  // 1. hash every position
  std::vector<uint32_t> hv(length);
  std::vector<uint8_t> hl(length);
  ParallelFor(length, threads, 1 << 16, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      uint32_t h, l;
      if (hf.Hash(seq + i, &h, &l)) {
        hv[i] = h;
        hl[i] = (uint8_t)l;
      } else {
        hl[i] = 0;
      }
    }
  });
  std::vector<KV> pairs;
  pairs.reserve(length / 2);
  uint32_t prev = 0;
  for (uint32_t i = 0; i < length; ++i) {
    if (hl[i]) {
      KV kv = {hv[i], i};
      if (prev != 0 && (prev - 1) >= hl[i])
        pairs.back() = kv;
      else
        pairs.push_back(kv);
      prev = hl[i];
    } else {
      prev = 0;
    }
  }
  std::vector<uint32_t>().swap(hv);
  std::vector<uint8_t>().swap(hl);
  RadixSortByHash(pairs, threads);
  // bucket boundaries
  std::vector<size_t> bstart;
  for (size_t i = 0; i < pairs.size(); ++i)
    if (i == 0 || pairs[i].hash != pairs[i - 1].hash) bstart.push_back(i);
  bstart.push_back(pairs.size());
  const size_t nb = bstart.size() - 1;

  const uint32_t L = common.subsequence_length;
  const uint32_t half = L / 2;
  const int maxmm = common.max_mismatches;
  const uint8_t *rmap = common.reduced_map;
  const bool fast = (L - half) <= 6 && half <= 7 && maxmm <= 1 && chunk.clustering;
  // the fast path also needs the hash digits to be the similarity codes
  bool same_alphabet = true;
  for (int c = 0; c < 25; ++c)
    if (hf.code_map[c] <= hf.max_code && hf.code_map[c] != rmap[c] + 1) same_alphabet = false;

  // 2-3. cluster per bucket, in parallel over contiguous bucket ranges
  const size_t ranges = std::max<size_t>(1, std::min<size_t>(nb, (size_t)threads * 8));
  std::vector<BucketOut> outs(ranges);
  std::atomic<size_t> next(0);
  auto work = [&]() {
    KeyTable table;
    std::vector<uint8_t> role;       // 0 value, 1 representative, 2 member
    std::vector<uint32_t> rep_of;    // member -> representative slot
    std::vector<uint32_t> members;   // per slot
    std::vector<uint32_t> keys;
    for (;;) {
      const size_t r = next.fetch_add(1);
      if (r >= ranges) break;
      const size_t b0 = nb * r / ranges, b1 = nb * (r + 1) / ranges;
      BucketOut &o = outs[r];
      for (size_t b = b0; b < b1; ++b) {
        const size_t s = bstart[b], e = bstart[b + 1], m = e - s;
        o.hashes.push_back(pairs[s].hash);
        role.assign(m, 0);
        rep_of.assign(m, 0);
        members.assign(m, 0);
        if (chunk.clustering && m > 1) {
          // windows touching the chunk start or a delimiter never cluster
          std::vector<bool> clean(m);
          for (size_t k = 0; k < m; ++k) {
            const uint32_t c = pairs[s + k].pos;
            bool ok = c >= half;
            if (ok)
              for (uint32_t t = 0; t < L; ++t)
                if (seq[c - half + t] == kDelimiter) {
                  ok = false;
                  break;
                }
            clean[k] = ok;
          }
          if (fast && same_alphabet) {
            table.Reset(m * half);
            keys.assign(m, 0);
            for (size_t k = 0; k < m; ++k) {
              if (!clean[k]) continue;
              const uint32_t c = pairs[s + k].pos;
              uint32_t key = 0;
              for (uint32_t t = 0; t < half; ++t) key = (key << 4) | rmap[seq[c - half + t]];
              uint32_t best = UINT32_MAX;
              for (uint32_t p = 0; p < half; ++p) {
                uint32_t v;
                if (table.Find(key | (0xFu << (4 * p)), &v)) best = std::min(best, v);
              }
              if (maxmm == 0) {
                // exact match only: a wildcard hit must also match exactly
                best = UINT32_MAX;
                uint32_t v;
                if (table.Find(key | (0xFu << 28), &v)) best = v;
              }
              if (best != UINT32_MAX) {
                role[k] = 2;
                rep_of[k] = best;
                ++members[best];
              } else {
                role[k] = 1;
                if (maxmm == 0) {
                  table.InsertIfAbsent(key | (0xFu << 28), (uint32_t)k);
                } else {
                  for (uint32_t p = 0; p < half; ++p)
                    table.InsertIfAbsent(key | (0xFu << (4 * p)), (uint32_t)k);
                }
              }
            }
          } else {
            // direct greedy (src/seed_searcher_database_parameters.cpp:151-178)
            std::vector<uint8_t> taken(m, 0);
            for (size_t k = 0; k < m; ++k) {
              if (!clean[k] || taken[k]) continue;
              taken[k] = 1;
              role[k] = 1;
              const uint32_t c = pairs[s + k].pos;
              for (size_t j = k + 1; j < m; ++j) {
                if (!clean[j] || taken[j]) continue;
                const uint32_t d = pairs[s + j].pos;
                int mm = 0;
                for (uint32_t t = 0; t < L; ++t)
                  mm += rmap[seq[c - half + t]] != rmap[seq[d - half + t]];
                if (mm <= maxmm) {
                  taken[j] = 1;
                  role[j] = 2;
                  rep_of[j] = (uint32_t)k;
                  ++members[k];
                }
              }
            }
          }
        }
        // emit: representatives with members, their members, the rest
        uint32_t nrep = 0, nval = 0;
        std::vector<uint32_t> slot(m, 0);
        for (size_t k = 0; k < m; ++k) {
          if (role[k] == 1 && members[k] > 0) {
            slot[k] = (uint32_t)o.mem_count.size();
            o.rep_pos.push_back(pairs[s + k].pos);
            o.mem_count.push_back(members[k]);
            ++nrep;
          } else if (role[k] != 2) {
            o.val_pos.push_back(pairs[s + k].pos);
            ++nval;
          }
        }
        if (nrep) {
          // members in position order, grouped by representative
          const size_t first_slot = o.mem_count.size() - nrep;
          std::vector<uint32_t> fill(nrep, 0);
          std::vector<uint32_t> base(nrep + 1, 0);
          for (uint32_t t = 0; t < nrep; ++t) base[t + 1] = base[t] + o.mem_count[first_slot + t];
          const size_t mem_base = o.mem_pos.size();
          o.mem_pos.resize(mem_base + base[nrep]);
          for (size_t k = 0; k < m; ++k) {
            if (role[k] == 2) {
              const uint32_t t = slot[rep_of[k]] - (uint32_t)first_slot;
              o.mem_pos[mem_base + base[t] + fill[t]++] = pairs[s + k].pos;
            }
          }
        }
        o.rep_count.push_back(nrep);
        o.val_count.push_back(nval);
      }
    }
  };
  {
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
  }
  std::vector<KV>().swap(pairs);
  // 4. concatenate the per-range outputs
  chunk.hashes.clear();
  chunk.rep_off.assign(1, 0);
  chunk.rep_pos.clear();
  chunk.mem_off.assign(1, 0);
  chunk.mem_pos.clear();
  chunk.val_off.assign(1, 0);
  chunk.val_pos.clear();
  for (size_t r = 0; r < ranges; ++r) {
    BucketOut &o = outs[r];
    chunk.hashes.insert(chunk.hashes.end(), o.hashes.begin(), o.hashes.end());
    for (size_t i = 0; i < o.rep_count.size(); ++i) {
      chunk.rep_off.push_back(chunk.rep_off.back() + o.rep_count[i]);
      chunk.val_off.push_back(chunk.val_off.back() + o.val_count[i]);
    }
    chunk.rep_pos.insert(chunk.rep_pos.end(), o.rep_pos.begin(), o.rep_pos.end());
    for (size_t i = 0; i < o.mem_count.size(); ++i)
      chunk.mem_off.push_back(chunk.mem_off.back() + o.mem_count[i]);
    chunk.mem_pos.insert(chunk.mem_pos.end(), o.mem_pos.begin(), o.mem_pos.end());
    chunk.val_pos.insert(chunk.val_pos.end(), o.val_pos.begin(), o.val_pos.end());
    BucketOut().hashes.swap(o.hashes);
    o = BucketOut();
  }
}

void BuildDatabase(const std::vector<FastaRecord> &subjects,
                   const DatabaseOptions &opt, const int *matrix, uint32_t nl,
                   Database *db) {
  // common parameters (src/aligner.cpp:114-168,
  // src/seed_searcher_database_parameters.cpp:41-78)
  uint32_t max_regular = 0;
  ReducedCodeMap(opt.alphabet_sets.c_str(), db->reduced_map, &max_regular);
  db->subsequence_length = opt.subsequence_length;
  const float rate = 0.9f;
  db->max_mismatches = (int)(opt.subsequence_length * (1 - rate));
  std::vector<int> scores(max_regular + 1, 0);
  for (uint32_t i = 0; i < nl && i < 25; ++i) {
    const uint8_t c = db->reduced_map[i];
    if (c <= max_regular) scores[c] = std::max(scores[c], matrix[i * nl + i]);
  }
  HashFunction &hf = db->hash;
  hf.max_code = max_regular + 1;
  hf.threshold = opt.seed_threshold;
  hf.max_length = 8;
  hf.shift = 0;
  for (uint32_t c = 1; c < hf.max_code; c <<= 1) ++hf.shift;
  hf.scores.assign(scores.size() + 1, 0);
  for (size_t i = 0; i < scores.size(); ++i) hf.scores[i + 1] = scores[i];
  for (int c = 0; c < 25; ++c) hf.code_map[c] = db->reduced_map[c] + 1;

  // chunks by residues (Database::ReadSequences, src/database.h:430-457)
  db->chunks.clear();
  db->total_length = 0;
  db->total_sequences = 0;
  db->max_length = 0;
  size_t i = 0;
  const uint8_t *table = ProteinEncodeTable();
  while (i < subjects.size()) {
    uint64_t sum = 0;
    size_t j = i;
    while (j < subjects.size()) {
      const uint64_t next_sum = sum + subjects[j].sequence.size();
      if (next_sum >= opt.chunk_size) break;
      sum = next_sum;
      ++j;
    }
    if (j == i) break;  // a subject longer than a chunk ends the build
    db->chunks.emplace_back();
    DbChunk &c = db->chunks.back();
    c.clustering = opt.clustering;
    uint64_t total = 1;
    for (size_t k = i; k < j; ++k) total += subjects[k].sequence.size() + 1;
    c.seq.resize(total);
    c.offsets.resize(j - i + 1);
    c.names.resize(j - i);
    c.seq[0] = kDelimiter;
    uint32_t off = 1;
    for (size_t k = i; k < j; ++k) {
      const std::string &s = subjects[k].sequence;
      c.offsets[k - i] = off;
      c.names[k - i] = subjects[k].name;
      for (size_t x = 0; x < s.size(); ++x) c.seq[off + x] = table[(uint8_t)s[x]];
      off += (uint32_t)s.size();
      c.seq[off++] = kDelimiter;
      db->max_length = std::max<uint32_t>(db->max_length, (uint32_t)s.size());
    }
    c.offsets[j - i] = off;
    db->total_length += sum;
    db->total_sequences += j - i;
    BuildChunkIndex(c, *db, opt);
    i = j;
  }
}

}  // namespace gz
