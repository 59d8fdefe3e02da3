// database.h -- the searched database: chunks of concatenated subjects plus
// the clustered seed index that `ghostz-gpu db` builds
// (src/database.h:440-514, src/database_chunk.h:188-210,
// src/seed_searcher_database_parameters.cpp:41-319).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "queries.h"

namespace gz {

// ReducedAlphabetVariableHashFunction state
// (src/reduced_alphabet_variable_hash_function.cpp:14-54): the code map is
// shifted by one so that 0 never appears in a hash digit.
struct HashFunction {
  uint8_t code_map[25];        // protein code -> reduced code + 1
  std::vector<int> scores;     // per shifted code
  uint32_t max_code = 0;       // largest shifted regular code
  int threshold = 39;
  uint32_t max_length = 8;
  uint32_t shift = 4;
  // CalculateHash; false when a non-regular code is met first
  inline bool Hash(const uint8_t *s, uint32_t *hash, uint32_t *len) const {
    uint32_t h = 0, l = 0;
    int score = 0;
    for (l = 0; l < max_length; ++l) {
      const uint8_t c0 = s[l];
      if (c0 < 25 && code_map[c0] <= max_code) {
        if (l >= 6 && score >= threshold) break;
        h = (h << shift) + code_map[c0];
        score += scores[code_map[c0]];
      } else {
        return false;
      }
    }
    *hash = h;
    *len = l;
    return true;
  }
};

struct DbChunk {
  std::vector<uint8_t> seq;       // leading delimiter, subject + delimiter ...
  std::vector<uint32_t> offsets;  // subject starts, number_sequences + 1
  std::vector<std::string> names;
  bool clustering = true;
  // seed index keyed by ascending hash
  std::vector<uint32_t> hashes;
  std::vector<uint32_t> rep_off, rep_pos;  // clusters with members, per hash
  std::vector<uint32_t> mem_off, mem_pos;  // members, per representative
  std::vector<uint32_t> val_off, val_pos;  // non-clustered positions, per hash
  uint32_t NumberSequences() const { return (uint32_t)names.size(); }
  // DatabaseChunk::GetId: subject containing a position
  uint32_t SubjectOf(uint32_t position) const;
};

struct DatabaseOptions {
  bool clustering = true;          // -C
  uint32_t subsequence_length = 10;  // -L
  int seed_threshold = 39;         // -s
  uint64_t chunk_size = 1ull << 30;  // -l (residues)
  std::string alphabet_sets = "A KR EDNQ C G H ILVM FYW P ST";
  int threads = 0;
};

struct Database {
  std::vector<DbChunk> chunks;
  uint64_t total_length = 0;     // residues (Database::database_length_)
  uint64_t total_sequences = 0;
  uint32_t max_length = 0;
  // common parameters (DatabaseChunk::GetSeedSearcherCommonParameters)
  uint32_t subsequence_length = 10;
  uint8_t reduced_map[25];       // similarity map
  HashFunction hash;
  int max_mismatches = 1;        // GetNumberMaxMismatchs
};

// Builds the database from subject records (chunked by residues).
void BuildDatabase(const std::vector<FastaRecord> &subjects,
                   const DatabaseOptions &opt, const int *matrix, uint32_t nl,
                   Database *db);

// Builds one chunk's seed index (exposed for tests).
void BuildChunkIndex(DbChunk &chunk, const Database &common,
                     const DatabaseOptions &opt);

}  // namespace gz
