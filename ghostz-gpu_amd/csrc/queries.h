// queries.h -- query preparation for the search path.
//
// Queries (src/queries.cpp), ProteinQuery (src/protein_query.cpp:22-60),
// TranslatedDnaQuery (src/translated_dna_query.{h,cpp}), Translator
// (src/translator.cpp:104-142) and FastaSequenceReader
// (src/fasta_sequence_reader.cpp:25-73).
#pragma once

#include <stdint.h>

#include <istream>
#include <string>
#include <vector>

#include "statistics.h"

namespace gz {

struct FastaRecord {
  std::string name;
  std::string sequence;  // upper-cased, whitespace removed
};

// Reads the next record; false at end of input.
bool ReadFasta(std::istream &in, FastaRecord *rec);

// Six frames: forward offsets 0,1,2 then reverse-complement offsets 0,1,2.
void TranslateSixFrames(const std::string &dna, std::vector<std::string> &frames);

// A prepared query chunk.  Query i's framed code sequence (leading and
// trailing delimiter) is codes[offsets[i] .. offsets[i+1]).
struct QuerySet {
  bool translated = false;
  std::vector<std::string> names;
  std::vector<uint8_t> codes;
  std::vector<uint64_t> offsets;
  std::vector<float> lambda, K;
  std::vector<uint32_t> dna_lengths;  // translated queries only
  uint32_t max_length = 0;

  size_t size() const { return names.size(); }
  uint32_t Length(size_t i) const { return (uint32_t)(offsets[i + 1] - offsets[i]); }
  const uint8_t *Sequence(size_t i) const { return codes.data() + offsets[i]; }
  // Query::GetRealSequenceLength and GetRealStart/GetRealEnd
  uint32_t RealLength(size_t i) const;
  uint32_t RealStart(size_t i, uint32_t pos) const;
  uint32_t RealEnd(size_t i, uint32_t pos) const;
  // distance of a framed position to the frame's delimiters
  // (Query::GetDistanceFromStartDelimiter / GetDistanceFromEndDelimiter)
  uint32_t DistanceFromStart(size_t i, uint32_t pos) const;
  uint32_t DistanceFromEnd(size_t i, uint32_t pos) const;
};

struct QueryOptions {
  bool translated = false;  // -q d -t p
  bool filter = true;       // -F T (SEG)
  int threads = 0;          // 0 = hardware concurrency
};

// Framed length a raw sequence of this length occupies (the chunk-size unit,
// Queries::GetSequenceLength).
uint32_t FramedLength(bool translated, uint32_t raw_length);

// Builds queries from raw records (names + residue/nucleotide strings).
void BuildQueries(const std::vector<FastaRecord> &records,
                  const QueryOptions &opt, const int *matrix, uint32_t nl,
                  QuerySet *out);

}  // namespace gz
