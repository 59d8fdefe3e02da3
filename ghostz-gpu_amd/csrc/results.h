// results.h -- host-side steps after the seed stage: chain filter, result
// merge, traceback re-extension and the BLAST-like tabular output.
#pragma once

#include <stdint.h>

#include <ostream>
#include <string>
#include <unordered_map>
#include <utility>
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

// One result build (traceback re-extension of both sides of a presearched
// result) and its outcome.
struct TracebackJob {
  uint32_t query_offset;  // concatenated position of the query's leading delimiter
  uint32_t query_length;  // framed length
  uint32_t hit_q, hit_d;  // query-local / chunk positions of the hit
};
constexpr uint32_t kTracebackStale = 1;     // read a byte the extension did not write
constexpr uint32_t kTracebackOverflow = 2;  // beyond the device scratch: redo on the host
struct TracebackOut {
  int score;
  uint32_t start_q, start_d, end_q, end_d;
  uint32_t length, mismatches, gap_openings, gap_extensions;
  uint32_t flags;
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

// The filter pass of ChainFilter on hits already in database-position order
// (the order std::sort gives; it is unique when no two hits share a
// database position).
void ChainFilterSorted(const uint8_t *query, uint32_t query_length, int cutoff,
                       const uint8_t *database, const int *matrix, uint32_t nl,
                       std::vector<Hit> &hits, std::vector<uint64_t> &chain_scratch);

// AlignerCommon::AddResults (src/aligner_common.h:147-274).
struct AlignmentPosition {
  uint32_t result_id;
  uint32_t position;
};
// The sorted start positions of the kept results per database subject.  The
// reference keeps a vector per subject of the chunk; one query holds at most
// max_results results, so a short list of the subjects in use (with a hash
// index once it grows) holds the same lists without touching a
// chunk-sized array per query.
class SubjectStarts {
 public:
  std::vector<AlignmentPosition> &Of(uint32_t subject);
  void Clear();

 private:
  std::vector<uint32_t> subjects_;
  std::vector<std::vector<AlignmentPosition> > lists_;
  size_t used_ = 0;
  std::unordered_map<uint32_t, uint32_t> index_;  // subject -> slot, when used_ > 16
};
// Per-thread scratch of AddResults, reused across calls.
struct AddResultsScratch {
  SubjectStarts starts;
  std::vector<Presearched> merged;
  std::vector<uint32_t> order;  // the new results by score, a prefix at a time
};
void AddResults(const DbChunk &chunk, int chunk_id, const SearchParameters &p,
                AddResultsScratch &scratch, std::vector<Presearched> &added,
                std::vector<Presearched> &results);

// Traceback re-extension of every presearched result of one chunk and the
// final per-query sort (AlignerBuildResultsThread, aligner_build_results_thread.h
// :64-190).  Runs on `threads` host threads; see results.cpp for how the
// result stays identical to the reference's one-thread order.
void BuildResults(const QuerySet &q, const Database &db, const SearchParameters &p,
                  int result_cutoff,
                  std::vector<std::vector<Presearched> > &presearched,
                  std::vector<std::vector<AlignmentResult> > &results, int threads,
                  uint64_t *sequential_chunks);

// Pieces of BuildResults for callers that build some results elsewhere:
// (query, result) pairs of chunk ci in the reference order; the host build of
// a chunk (parallel, replayed in the reference order when a traceback read a
// byte left by an earlier extension); the sequential replay itself; one
// result with a fresh extender (false when it read such a byte); the final
// per-query order.
std::vector<std::pair<uint32_t, uint32_t> > ChunkResultIds(
    const std::vector<std::vector<Presearched> > &pre, size_t ci);
// the same for queries [qb, qe) only (a contiguous part of the above)
std::vector<std::pair<uint32_t, uint32_t> > ChunkResultIds(
    const std::vector<std::vector<Presearched> > &pre, size_t ci, size_t qb, size_t qe);
void BuildChunkHost(const QuerySet &q, const DbChunk &c, const SearchParameters &p, int cutoff,
                    std::vector<std::vector<Presearched> > &pre,
                    std::vector<std::vector<AlignmentResult> > &results,
                    const std::vector<std::pair<uint32_t, uint32_t> > &ids, int threads,
                    uint64_t *sequential_chunks);
void ReplayChunkHost(const QuerySet &q, const DbChunk &c, const SearchParameters &p, int cutoff,
                     std::vector<std::vector<Presearched> > &pre,
                     std::vector<std::vector<AlignmentResult> > &results,
                     const std::vector<std::pair<uint32_t, uint32_t> > &ids);
bool BuildOneFresh(const QuerySet &q, uint32_t qi, const DbChunk &c, const Presearched &pr,
                   const SearchParameters &p, int cutoff, AlignmentResult *out);
// The results of chunk c at positions `flagged` of `ids` (the chunk's ids
// in ChunkResultIds order) whose traceback read op bytes that the chunk's
// earlier extensions left, rebuilt with those bytes: ReplayChunkHost's
// result for them, without replaying the rest.  False when a read falls
// outside what it models (the caller then replays the chunk).
bool ResolveStaleChunk(const QuerySet &q, const DbChunk &c, const SearchParameters &p,
                       int cutoff, std::vector<std::vector<Presearched> > &pre,
                       std::vector<std::vector<AlignmentResult> > &results,
                       const std::vector<std::pair<uint32_t, uint32_t> > &ids,
                       const std::vector<size_t> &flagged);
void SortResults(std::vector<std::vector<AlignmentResult> > &results);
void SortResults(std::vector<std::vector<AlignmentResult> > &results, size_t qb, size_t qe);
// The result build of one hit with a fresh extender on the framed query
// (query[0] is its leading delimiter) and the chunk codes.
TracebackOut TracebackOneHost(const uint8_t *query, uint32_t query_length,
                              const uint8_t *database, uint32_t hit_q, uint32_t hit_d,
                              const SearchParameters &p, int cutoff);

// A double as printf's %g writes it (six significant digits, rounded half
// to even on the exact binary value, trailing zeros dropped), which is how
// the reference's std::ostream writes the identity, E-value and bit score
// (src/aligner_common.h:372-389).  Writes at most 13 characters, so a
// 32-byte buffer always holds the text; returns its end.
char *FormatG6(double v, char *p, char *end);

// AlignerCommon::WriteOutput (src/aligner_common.h:345-397); the text form
// appends the lines of queries [qb, qe) to *out, formatted on up to
// `threads` threads (0: all cores).
void WriteOutput(std::ostream &os, const QuerySet &q, const Database &db,
                 const std::vector<std::vector<AlignmentResult> > &results);
void WriteOutputText(std::string *out, const QuerySet &q, const Database &db,
                     const std::vector<std::vector<AlignmentResult> > &results, size_t qb,
                     size_t qe, int threads);

}  // namespace gz
