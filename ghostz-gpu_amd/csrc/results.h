// results.h -- host-side steps after the seed stage: chain filter, result
// merge, traceback re-extension and the BLAST-like tabular output.
#pragma once

#include <stdint.h>

#include <ostream>
#include <string>
#include <vector>

#include "database.h"
#include "queries.h"
#include "statistics.h"

namespace gz {

struct Coordinate {
  uint32_t q;
  uint32_t d;
};

struct Hit {
  uint32_t q;  // query-local position
  uint32_t d;  // chunk position
};

// AlignerCommon::PresearchedResult (src/aligner_common.h:70-78)
struct Presearched {
  int chunk;
  uint32_t subject;
  int score;
  Coordinate hit, start, end;
  uint32_t hit_count;
};

// AlignerCommon::Result (src/aligner_common.h:85-98)
struct AlignmentResult {
  uint32_t chunk;
  uint32_t subject;
  int score;
  uint32_t length, mismatches, gap_openings, gap_extensions;
  Coordinate hit, start, end;
};

struct SearchParameters {
  const int *matrix = nullptr;  // host nl x nl
  uint32_t nl = 25;
  int gap_open = 11;
  int gap_extension = 1;
  float normalized_ungapped_cutoff = 7.0f;
  float normalized_gapped_trigger = 22.0f;
  float normalized_presearch_cutoff = 15.0f;
  float normalized_result_cutoff = 25.0f;
  uint32_t max_results = 10;              // -b
  uint32_t max_results_per_subject = 1;   // -v
};

// ChainFilter::Filter (src/chain_filter.cpp:24-99); `hits` must be in the
// presearch enumeration order (the sort below is not stable).
void ChainFilter(const uint8_t *query, uint32_t query_length, int cutoff,
                 const uint8_t *database, const int *matrix, uint32_t nl,
                 std::vector<Hit> &hits, std::vector<uint64_t> &chain_scratch);

// AlignerCommon::AddResults (src/aligner_common.h:147-274).
struct AlignmentPosition {
  uint32_t result_id;
  uint32_t position;
};
void AddResults(const DbChunk &chunk, int chunk_id, const SearchParameters &p,
                std::vector<std::vector<AlignmentPosition> > &starts,
                std::vector<Presearched> &added, std::vector<Presearched> &results);

// Traceback re-extension of every presearched result of one chunk and the
// final per-query sort (AlignerBuildResultsThread, aligner_build_results_thread.h
// :64-190).  Runs on `threads` host threads; see results.cpp for how the
// result stays identical to the reference's one-thread order.
void BuildResults(const QuerySet &q, const Database &db, const SearchParameters &p,
                  int result_cutoff,
                  std::vector<std::vector<Presearched> > &presearched,
                  std::vector<std::vector<AlignmentResult> > &results, int threads,
                  uint64_t *sequential_chunks);

// AlignerCommon::WriteOutput (src/aligner_common.h:345-397).
void WriteOutput(std::ostream &os, const QuerySet &q, const Database &db,
                 const std::vector<std::vector<AlignmentResult> > &results);

}  // namespace gz
