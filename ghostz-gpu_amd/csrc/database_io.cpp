// database_io.cpp -- see database_io.h.
#include "database_io.h"

#include <stdio.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>

#include "gz_common.h"

namespace gz {

namespace {

std::string ChunkPrefix(const std::string &prefix, size_t i) {
  std::ostringstream s;
  s << prefix << "_" << i;  // Database::Save / SetChunk (src/database.h:245, :385)
  return s.str();
}

// Binary input with one error message for the first short read.
struct In {
  std::ifstream f;
  std::string path;
  std::string *err;
  bool ok = true;
  In(const std::string &p, std::string *e) : f(p.c_str(), std::ios::binary), path(p), err(e) {
    if (!f) Fail("cannot open");
  }
  void Fail(const std::string &why) {
    if (ok && err) *err = path + ": " + why;
    ok = false;
  }
  template <typename T>
  T Get() {
    T v{};
    if (ok && !f.read((char *)&v, sizeof(v))) Fail("truncated");
    return v;
  }
  template <typename T>
  void GetArray(std::vector<T> *v, size_t n) {
    if (!ok) return;
    if (n > (size_t)1 << 40) {
      Fail("implausible array length");
      return;
    }
    v->resize(n);
    if (n && !f.read((char *)v->data(), (std::streamsize)(n * sizeof(T)))) Fail("truncated");
  }
  // a std::vector saved as (size_t count, elements)
  template <typename T>
  void GetVector(std::vector<T> *v) {
    GetArray(v, Get<uint64_t>());
  }
};

struct Out {
  std::ofstream f;
  Out(const std::string &p) : f(p.c_str(), std::ios::binary) {}
  template <typename T>
  void Put(const T &v) {
    f.write((const char *)&v, sizeof(v));
  }
  template <typename T>
  void PutVector(const T *p, size_t n) {
    Put<uint64_t>(n);
    if (n) f.write((const char *)p, (std::streamsize)(n * sizeof(*p)));
  }
};

// ReducedAlphabetVariableHashFunction::Save / Load field by field
// (src/reduced_alphabet_variable_hash_function.cpp:133-171).
struct SavedHash {
  uint8_t can_calculate = 0;
  uint8_t max_code = 0;  // largest regular code + 1
  int score_threshold = 0;
  uint32_t max_length = 0, shift = 0, max_hash = 0;
  std::vector<uint8_t> code_map;  // shifted by one
  std::vector<int> code_scores;   // per shifted code, [0] = 0
};

void ReadHash(In &in, SavedHash *h) {
  h->can_calculate = in.Get<uint8_t>();
  h->max_code = in.Get<uint8_t>();
  h->score_threshold = in.Get<int>();
  h->max_length = in.Get<uint32_t>();
  h->shift = in.Get<uint32_t>();
  h->max_hash = in.Get<uint32_t>();
  in.GetVector(&h->code_map);
  in.GetVector(&h->code_scores);
}

SavedHash FromHashFunction(const HashFunction &hf) {
  SavedHash h;
  h.max_code = (uint8_t)hf.max_code;
  h.score_threshold = hf.threshold;
  h.max_length = hf.max_length;
  h.shift = hf.shift;
  // SetShiftSize / SetMaxHash (:106-131)
  const uint32_t bits = h.shift * h.max_length;
  const bool fits = !(bits > 32 || (bits == 32 && (uint32_t)h.max_code + 1 == (1u << h.shift)));
  h.can_calculate = fits ? 1 : 0;
  if (fits)
    for (uint32_t i = 0; i < h.max_length; ++i) h.max_hash = (h.max_hash << h.shift) + h.max_code;
  h.code_map.assign(hf.code_map, hf.code_map + kNumberLetters);
  h.code_scores = hf.scores;
  return h;
}

void WriteHash(Out &o, const SavedHash &h) {
  o.Put(h.can_calculate);
  o.Put(h.max_code);
  o.Put(h.score_threshold);
  o.Put(h.max_length);
  o.Put(h.shift);
  o.Put(h.max_hash);
  o.PutVector(h.code_map.data(), h.code_map.size());
  o.PutVector(h.code_scores.data(), h.code_scores.size());
}

// DenseHashInvertedIndex::Load (src/dense_hash_inverted_index.h:123-144):
// (hash -> slot) pairs, slot offsets, values.
template <typename V>
struct DenseIndex {
  std::map<uint32_t, uint32_t> slot;
  std::vector<uint32_t> offsets;
  std::vector<V> values;
};

template <typename V>
void ReadDense(In &in, DenseIndex<V> *d) {
  const uint64_t n = in.Get<uint64_t>();
  if (n > (uint64_t)1 << 32) in.Fail("implausible index size");
  for (uint64_t i = 0; i < n && in.ok; ++i) {
    const uint32_t k = in.Get<uint32_t>();
    const uint32_t v = in.Get<uint32_t>();
    d->slot[k] = v;
  }
  in.GetVector(&d->offsets);
  in.GetVector(&d->values);
  if (!in.ok) return;
  for (size_t i = 1; i < d->offsets.size(); ++i)
    if (d->offsets[i] < d->offsets[i - 1]) in.Fail("index offsets decrease");
  if (!d->offsets.empty() && d->offsets.back() > d->values.size())
    in.Fail("index offsets past the values");
  for (const auto &kv : d->slot)
    if ((size_t)kv.second + 1 >= d->offsets.size()) in.Fail("index slot out of range");
}

// PerfectHashInvertedIndex::Load (src/perfect_hash_inverted_index.h:112-124).
struct PerfectIndex {
  uint32_t max_key = 0;
  std::vector<uint32_t> offsets, values;
};

void ReadPerfect(In &in, PerfectIndex *p) {
  p->max_key = in.Get<uint32_t>();
  in.GetVector(&p->offsets);
  in.GetVector(&p->values);
  if (!in.ok) return;
  for (size_t i = 1; i < p->offsets.size(); ++i)
    if (p->offsets[i] < p->offsets[i - 1]) in.Fail("member offsets decrease");
  if (!p->offsets.empty() && p->offsets.back() > p->values.size())
    in.Fail("member offsets past the values");
}

// The seed index stream of SeedSearcherDatabaseParameters::Save
// (src/seed_searcher_database_parameters.cpp:363-376) into the chunk's
// index arrays: per hash (ascending) the representatives in stored order,
// each with its cluster's members, and the non-clustered values in stored
// order.
bool ReadIndex(const std::string &path, DbChunk *c, uint32_t *subsequence_length,
               float *cluster_rate, std::string *err) {
  In in(path, err);
  const uint8_t clustering = in.Get<uint8_t>();
  const uint8_t delimiter = in.Get<uint8_t>();
  *subsequence_length = in.Get<uint32_t>();
  SavedHash h;
  ReadHash(in, &h);
  DenseIndex<uint32_t> values;
  ReadDense(in, &values);
  ReadHash(in, &h);
  DenseIndex<std::pair<uint32_t, uint32_t> > reps;  // (cluster id, position)
  ReadDense(in, &reps);
  PerfectIndex members;
  ReadPerfect(in, &members);
  *cluster_rate = in.Get<float>();
  std::vector<uint8_t> map;
  in.GetVector(&map);
  if (!in.ok) return false;
  if (delimiter != kDelimiter) {
    in.Fail("sequence delimiter is not the protein delimiter");
    return false;
  }
  c->clustering = clustering != 0;
  std::vector<uint32_t> all;
  for (const auto &kv : values.slot) all.push_back(kv.first);
  for (const auto &kv : reps.slot) all.push_back(kv.first);
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  c->hashes = all;
  c->rep_off.assign(1, 0);
  c->rep_pos.clear();
  c->mem_off.assign(1, 0);
  c->mem_pos.clear();
  c->val_off.assign(1, 0);
  c->val_pos.clear();
  for (uint32_t hash : all) {
    auto r = reps.slot.find(hash);
    if (r != reps.slot.end()) {
      for (uint32_t k = reps.offsets[r->second]; k < reps.offsets[r->second + 1]; ++k) {
        c->rep_pos.push_back(reps.values[k].second);
        const uint32_t cid = reps.values[k].first;
        if ((size_t)cid + 1 < members.offsets.size())
          for (uint32_t m = members.offsets[cid]; m < members.offsets[cid + 1]; ++m)
            c->mem_pos.push_back(members.values[m]);
        c->mem_off.push_back((uint32_t)c->mem_pos.size());
      }
    }
    c->rep_off.push_back((uint32_t)c->rep_pos.size());
    auto v = values.slot.find(hash);
    if (v != values.slot.end())
      for (uint32_t k = values.offsets[v->second]; k < values.offsets[v->second + 1]; ++k)
        c->val_pos.push_back(values.values[k]);
    c->val_off.push_back((uint32_t)c->val_pos.size());
  }
  return true;
}

// SeedSearcherDatabaseParameters::BuildCommonParameters(istream)
// (src/seed_searcher_database_parameters.cpp:347-361).
bool ReadCommon(const std::string &path, Database *db, std::string *err) {
  In in(path, err);
  const uint8_t delimiter = in.Get<uint8_t>();
  const uint32_t L = in.Get<uint32_t>();
  std::vector<uint8_t> map;
  in.GetVector(&map);
  SavedHash h;
  ReadHash(in, &h);
  if (!in.ok) return false;
  if (delimiter != kDelimiter) {
    in.Fail("sequence delimiter is not the protein delimiter");
    return false;
  }
  if (map.size() < kNumberLetters || h.code_map.size() < kNumberLetters) {
    in.Fail("code maps shorter than the protein alphabet");
    return false;
  }
  db->subsequence_length = L;
  for (uint32_t i = 0; i < kNumberLetters; ++i) {
    db->reduced_map[i] = map[i];
    db->hash.code_map[i] = h.code_map[i];
  }
  db->hash.scores = h.code_scores;
  db->hash.max_code = h.max_code;
  db->hash.threshold = h.score_threshold;
  db->hash.max_length = h.max_length;
  db->hash.shift = h.shift;
  if (db->hash.max_length > 8 || h.code_scores.size() <= h.max_code) {
    in.Fail("hash function outside what the seed search supports");
    return false;
  }
  return true;
}

}  // namespace

bool LoadDatabase(const std::string &prefix, Database *db, std::string *err) {
  // Database::LoadInfomation (src/database.h:517-538)
  In in(prefix + ".inf", err);
  const int32_t number_chunks = in.Get<int32_t>();
  const uint32_t max_length = in.Get<uint32_t>();
  const uint64_t total_length = in.Get<uint64_t>();
  const uint64_t total_sequences = in.Get<uint64_t>();
  const uint8_t delimiter = in.Get<uint8_t>();
  if (!in.ok) return false;
  if (number_chunks < 0 || delimiter != kDelimiter) {
    in.Fail("not a protein database");
    return false;
  }
  db->chunks.assign((size_t)number_chunks, DbChunk());
  db->max_length = max_length;
  db->total_length = total_length;
  db->total_sequences = total_sequences;
  for (int32_t i = 0; i < number_chunks; ++i) {
    const std::string p = ChunkPrefix(prefix, (size_t)i);
    DbChunk &c = db->chunks[(size_t)i];
    // DatabaseChunk::LoadInfomation / LoadOffsets / LoadNames /
    // LoadConcatenatedSequence (src/database_chunk.h:303-370)
    In ci(p + ".inf", err);
    const uint32_t n = ci.Get<uint32_t>();
    const uint32_t length = ci.Get<uint32_t>();
    if (!ci.ok) return false;
    In co(p + ".off", err);
    co.GetArray(&c.offsets, (size_t)n + 1);
    if (!co.ok) return false;
    In cs(p + ".seq", err);
    cs.GetArray(&c.seq, length);
    if (!cs.ok) return false;
    std::ifstream names((p + ".nam").c_str());
    if (!names) {
      if (err) *err = p + ".nam: cannot open";
      return false;
    }
    c.names.resize(n);
    for (uint32_t k = 0; k < n; ++k)
      if (!std::getline(names, c.names[k])) {
        if (err) *err = p + ".nam: fewer names than sequences";
        return false;
      }
    // the layout the search relies on: a leading delimiter, subjects
    // separated by delimiters, offsets at subject starts
    bool layout = length >= 1 && c.seq[0] == kDelimiter && c.offsets[n] == length;
    for (uint32_t k = 0; layout && k < n; ++k)
      layout = c.offsets[k] >= 1 && c.offsets[k] < c.offsets[k + 1] &&
               c.seq[c.offsets[k] - 1] == kDelimiter && c.seq[c.offsets[k + 1] - 1] == kDelimiter;
    for (uint32_t x = 0; layout && x < length; ++x) layout = c.seq[x] <= kDelimiter;
    if (!layout) {
      if (err) *err = p + ": chunk sequence and offsets do not match";
      return false;
    }
    if (i == 0 && !ReadCommon(p + ".scp", db, err)) return false;
    uint32_t L = 0;
    float rate = 0.f;
    if (!ReadIndex(p + ".sdp", &c, &L, &rate, err)) return false;
    if (L != db->subsequence_length) {
      if (err) *err = p + ".sdp: subsequence length differs from the common parameters";
      return false;
    }
    // SeedSearcherDatabaseParameters::GetNumberMaxMismatchs (:402-404)
    db->max_mismatches = (int)(L * (1 - rate));
    for (uint32_t pos : c.rep_pos)
      if (pos >= length) layout = false;
    for (uint32_t pos : c.mem_pos)
      if (pos >= length) layout = false;
    for (uint32_t pos : c.val_pos)
      if (pos >= length) layout = false;
    if (!layout) {
      if (err) *err = p + ".sdp: seed positions past the chunk";
      return false;
    }
  }
  return true;
}

bool SaveDatabase(const Database &db, const std::string &prefix, std::string *err) {
  const SavedHash h = FromHashFunction(db.hash);
  const float rate = 0.9f;  // SeedSearcherDatabaseParameters::Build (:48)
  for (size_t i = 0; i < db.chunks.size(); ++i) {
    const DbChunk &c = db.chunks[i];
    const std::string p = ChunkPrefix(prefix, i);
    const uint32_t n = c.NumberSequences(), length = (uint32_t)c.seq.size();
    {
      Out o(p + ".inf");
      o.Put(n);
      o.Put(length);
    }
    {
      Out o(p + ".off");
      o.f.write((const char *)c.offsets.data(), (std::streamsize)(c.offsets.size() * 4));
    }
    {
      std::ofstream o((p + ".nam").c_str());
      for (const std::string &s : c.names) o << s << "\n";
    }
    {
      Out o(p + ".seq");
      o.f.write((const char *)c.seq.data(), (std::streamsize)c.seq.size());
    }
    {
      // SaveCommonParameters (:378-389)
      Out o(p + ".scp");
      o.Put<uint8_t>(kDelimiter);
      o.Put<uint32_t>(db.subsequence_length);
      o.PutVector(db.reduced_map, kNumberLetters);
      WriteHash(o, h);
    }
    {
      // Save (:363-376): slots in ascending hash order, cluster ids in
      // representative order
      Out o(p + ".sdp");
      o.Put<uint8_t>(c.clustering ? 1 : 0);
      o.Put<uint8_t>(kDelimiter);
      o.Put<uint32_t>(db.subsequence_length);
      WriteHash(o, h);
      std::vector<uint32_t> voff(1, 0);
      std::vector<std::pair<uint32_t, uint32_t> > vkeys;
      for (size_t k = 0; k < c.hashes.size(); ++k)
        if (c.val_off[k + 1] > c.val_off[k]) {
          vkeys.emplace_back(c.hashes[k], (uint32_t)vkeys.size());
          voff.push_back(c.val_off[k + 1]);
        }
      o.Put<uint64_t>(vkeys.size());
      for (const auto &kv : vkeys) {
        o.Put(kv.first);
        o.Put(kv.second);
      }
      // the values, concatenated in slot order, are val_pos itself
      o.PutVector(voff.data(), voff.size());
      o.PutVector(c.val_pos.data(), c.val_pos.size());
      WriteHash(o, h);
      std::vector<uint32_t> roff(1, 0);
      std::vector<std::pair<uint32_t, uint32_t> > rkeys, rvals;
      for (size_t k = 0; k < c.hashes.size(); ++k)
        if (c.rep_off[k + 1] > c.rep_off[k]) {
          rkeys.emplace_back(c.hashes[k], (uint32_t)rkeys.size());
          for (uint32_t r = c.rep_off[k]; r < c.rep_off[k + 1]; ++r)
            rvals.emplace_back(r, c.rep_pos[r]);
          roff.push_back((uint32_t)rvals.size());
        }
      o.Put<uint64_t>(rkeys.size());
      for (const auto &kv : rkeys) {
        o.Put(kv.first);
        o.Put(kv.second);
      }
      o.PutVector(roff.data(), roff.size());
      o.PutVector(rvals.data(), rvals.size());
      const uint32_t reps = (uint32_t)c.rep_pos.size();
      o.Put<uint32_t>(reps ? reps - 1 : 0);
      o.PutVector(c.mem_off.data(), c.mem_off.size());
      o.PutVector(c.mem_pos.data(), c.mem_pos.size());
      o.Put(rate);
      o.PutVector(db.reduced_map, kNumberLetters);
      if (!o.f) {
        if (err) *err = p + ".sdp: write failed";
        return false;
      }
    }
  }
  // Database::SaveInformation (src/database.h:540-554)
  Out o(prefix + ".inf");
  o.Put<int32_t>((int32_t)db.chunks.size());
  o.Put<uint32_t>(db.max_length);
  o.Put<uint64_t>(db.total_length);
  o.Put<uint64_t>(db.total_sequences);
  o.Put<uint8_t>(kDelimiter);
  if (!o.f) {
    if (err) *err = prefix + ".inf: write failed";
    return false;
  }
  return true;
}

}  // namespace gz
