// queries.cpp -- see queries.h.
#include "queries.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstring>
#include <thread>

#include "alphabet.h"
#include "gz_common.h"
#include "seg.h"

namespace gz {

bool ReadFasta(std::istream &in, FastaRecord *rec) {
  rec->name.clear();
  rec->sequence.clear();
  std::string line;
  // skip to the next header line
  while (in.good() && (line.empty() || line[0] != '>')) {
    if (!std::getline(in, line)) return false;
  }
  if (line.empty() || line[0] != '>') return false;
  if (!line.empty() && line.back() == '\r') line.pop_back();
  const size_t p = line.find_first_not_of("> ");
  if (p != std::string::npos) rec->name = line.substr(p);
  // sequence lines until the next header (left in the stream)
  for (;;) {
    const int c = in.peek();
    if (c == EOF || c == '>') break;
    if (!std::getline(in, line)) break;
    size_t a = 0, b = line.size();
    while (a < b && std::isspace((unsigned char)line[a])) ++a;
    while (b > a && std::isspace((unsigned char)line[b - 1])) --b;
    rec->sequence.append(line, a, b - a);
  }
  for (size_t i = 0; i < rec->sequence.size(); ++i)
    rec->sequence[i] = (char)std::toupper((unsigned char)rec->sequence[i]);
  return true;
}

namespace {

// Standard genetic code, codon -> amino acid ('*' = stop), indexed by
// (base0 * 16 + base1 * 4 + base2) with A=0 C=1 G=2 T=3.
const char kCodonTable[] =
    "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";

inline int BaseIndex(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
  }
}

inline char Complement(char c) {
  switch (c) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    default: return 'N';
  }
}

}  // namespace

void TranslateSixFrames(const std::string &dna, std::vector<std::string> &frames) {
  frames.clear();
  const size_t n = dna.size();
  if (n < 3) return;
  std::string rc(dna.rbegin(), dna.rend());
  for (size_t i = 0; i < n; ++i) rc[i] = Complement(rc[i]);
  const std::string *strands[2] = {&dna, &rc};
  for (int s = 0; s < 2; ++s) {
    const std::string &seq = *strands[s];
    for (size_t off = 0; off < 3; ++off) {
      std::string p;
      for (size_t j = off; j + 2 < n; j += 3) {
        const int a = BaseIndex(seq[j]), b = BaseIndex(seq[j + 1]),
                  c = BaseIndex(seq[j + 2]);
        p.push_back((a < 0 || b < 0 || c < 0) ? 'X' : kCodonTable[a * 16 + b * 4 + c]);
      }
      frames.push_back(p);
    }
  }
}

uint32_t FramedLength(bool translated, uint32_t len) {
  if (translated) return len < 3 ? 0 : 7 + (len / 3) * 6;
  return len == 0 ? 0 : len + 2;
}

uint32_t QuerySet::RealLength(size_t i) const {
  return translated ? dna_lengths[i] / 3 : Length(i) - 2;
}

uint32_t QuerySet::RealStart(size_t i, uint32_t pos) const {
  if (!translated) return pos;
  const uint32_t dl = dna_lengths[i];
  const uint32_t base = dl / 3 + 1;
  const uint32_t f = (pos - 1) / base;
  if (f < 3) return ((pos % base) - 1) * 3 + f + 1;
  return dl - (((pos % base) - 1) * 3 + f - 2) + 1;
}

uint32_t QuerySet::RealEnd(size_t i, uint32_t pos) const {
  if (!translated) return pos;
  const uint32_t base = dna_lengths[i] / 3 + 1;
  const uint32_t f = (pos - 1) / base;
  return f < 3 ? RealStart(i, pos) + 2 : RealStart(i, pos) - 2;
}

uint32_t QuerySet::DistanceFromStart(size_t i, uint32_t pos) const {
  if (!translated) return pos;
  const uint32_t base = dna_lengths[i] / 3 + 1;
  const uint32_t f = (pos - 1) / base;
  return pos - f * base;
}

uint32_t QuerySet::DistanceFromEnd(size_t i, uint32_t pos) const {
  if (!translated) return Length(i) - 1 - pos;
  const uint32_t base = dna_lengths[i] / 3 + 1;
  const uint32_t f = (pos - 1) / base;
  return (f + 1) * base - pos;
}

namespace {

// One query's framed codes (ProteinQuery::SetEncodedSequence /
// TranslatedDnaQuery::SetEncodedSequence).
void FrameQuery(const std::string &raw, bool translated, bool filter,
                const Seg &seg, std::vector<uint8_t> &out) {
  const uint8_t *table = ProteinEncodeTable();
  out.clear();
  if (!translated) {
    out.assign(raw.size() + 2, 0);
    out[0] = kDelimiter;
    std::string masked;
    int status = 0;
    if (filter) status = seg.Mask(raw, &masked); else masked = raw;
    if (status == 0) {
      for (size_t k = 0; k < masked.size(); ++k) out[1 + k] = table[(uint8_t)masked[k]];
      out.back() = kDelimiter;
    }
    return;
  }
  std::vector<std::string> frames;
  TranslateSixFrames(raw, frames);
  out.assign(FramedLength(true, (uint32_t)raw.size()), 0);
  if (out.empty()) return;
  out[0] = kDelimiter;
  size_t off = 1;
  const size_t base = raw.size() / 3;
  for (int f = 0; f < 6; ++f) {
    const std::string &s = frames[f];
    std::string masked;
    int status = 0;
    if (filter) status = seg.Mask(s, &masked); else masked = s;
    if (status == 0) {
      for (size_t k = 0; k < s.size(); ++k) out[off + k] = table[(uint8_t)masked[k]];
    } else {
      for (size_t k = 0; k < s.size(); ++k) out[off + k] = kUnknownCode;
    }
    off += s.size();
    if (s.size() < base) out[off++] = kDelimiter;
    out[off++] = kDelimiter;
  }
}

}  // namespace

void BuildQueries(const std::vector<FastaRecord> &records,
                  const QueryOptions &opt, const int *matrix, uint32_t nl,
                  QuerySet *out) {
  const size_t n = records.size();
  out->translated = opt.translated;
  out->names.resize(n);
  out->lambda.assign(n, 0.f);
  out->K.assign(n, 0.f);
  out->dna_lengths.assign(opt.translated ? n : 0, 0);
  std::vector<std::vector<uint8_t> > framed(n);
  Statistics stats;
  KarlinParameters ideal;
  stats.IdealUngapped(matrix, nl, &ideal);
  int threads = opt.threads > 0 ? opt.threads : (int)std::thread::hardware_concurrency();
  threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(1, n / 16)));
  std::atomic<size_t> next(0);
  auto work = [&]() {
    Seg seg(12);
    for (;;) {
      const size_t begin = next.fetch_add(64);
      if (begin >= n) break;
      const size_t end = std::min(n, begin + 64);
      for (size_t i = begin; i < end; ++i) {
        out->names[i] = records[i].name;
        FrameQuery(records[i].sequence, opt.translated, opt.filter, seg, framed[i]);
        KarlinParameters k;
        if (framed[i].empty() ||
            !stats.QueryUngapped(framed[i].data(), (uint32_t)framed[i].size(),
                                 matrix, nl, &k))
          k = ideal;
        out->lambda[i] = k.lambda;
        out->K[i] = k.K;
        if (opt.translated) out->dna_lengths[i] = (uint32_t)records[i].sequence.size();
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  out->offsets.assign(n + 1, 0);
  for (size_t i = 0; i < n; ++i) out->offsets[i + 1] = out->offsets[i] + framed[i].size();
  out->codes.resize(out->offsets[n]);
  out->max_length = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!framed[i].empty())
      memcpy(out->codes.data() + out->offsets[i], framed[i].data(), framed[i].size());
    out->max_length = std::max<uint32_t>(out->max_length, (uint32_t)framed[i].size());
  }
}

}  // namespace gz
