/*
 * gz_oracle.h -- C ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of the
 * GHOSTZ-GPU reference CPU aligner (`aln -g 0`, one presearch thread).  It is
 * the checker for the HIP product path: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product library never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Positions are
 * indexes into delimiter-framed code arrays exactly as in the reference.
 */
#ifndef GZ_ORACLE_H_
#define GZ_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- leaf algorithms ---------------------------------------------------- */

/* UngappedExtender::ExtendOneSideScoreOnly (src/ungapped_extender.h:34-82).
 * seq0 = query side, seq1 = database side, both already offset to the first
 * compared code; matrix is row-major [seq1 code][seq0 code]. */
int gzo_ungapped_one_side(const uint8_t *seq0, const uint8_t *seq1,
                          uint8_t delimiter, int reversed, const int *matrix,
                          uint32_t number_letters, int cutoff);

/* UngappedExtender::ExtendOneSide (src/ungapped_extender.h:84-143):
 * score plus best positions (relative to the start). */
int gzo_ungapped_one_side_positions(const uint8_t *seq0, const uint8_t *seq1,
                                    uint8_t delimiter, int reversed,
                                    const int *matrix, uint32_t number_letters,
                                    int cutoff, int *best_seq0_position,
                                    int *best_seq1_position);

/* SeedSearcher::UngappedExtend (src/seed_searcher.h:133-158): reverse from
 * (qpos-1, dpos-1); forward from (qpos, dpos) only while sum <= trigger. */
int gzo_ungapped_with_trigger(const uint8_t *query, uint32_t query_position,
                              const uint8_t *database,
                              uint32_t database_position, uint8_t delimiter,
                              const int *matrix, uint32_t number_letters,
                              int cutoff, int trigger);

/* DistanceCalculator::CalculateDistance (src/distance_calculator.cpp:35-54). */
int gzo_distance(const uint8_t *query, uint32_t query_center,
                 const uint8_t *database, uint32_t database_center,
                 uint32_t subsequence_length, const uint8_t *reduced_code_map);

/* GappedExtender::ExtendOneSide (src/gapped_extender.cpp:70-304, score-only
 * variant :306-471).  When edits != NULL the traceback is produced
 * (PickOptimalPath :473-531) as one op per aligned column, in
 * EditBlocks order (0 = gap in seq0, 1 = substitution, 2 = gap in seq1);
 * *number_edits receives the count (edits must hold seq0_len+seq1_len+2).
 * *dp_cells (optional) receives the number of DP cells evaluated in the
 * inner loop (:410-443), the cost the learned scheduler predicts. */
int gzo_gapped_one_side(const uint8_t *seq0, uint32_t seq0_length,
                        const uint8_t *seq1, uint8_t delimiter, int reversed,
                        const int *matrix, uint32_t number_letters,
                        int gap_open, int gap_extension, int cutoff,
                        int *best_score, int *best_seq0_position,
                        int *best_seq1_position, uint8_t *edits,
                        uint32_t *number_edits, uint64_t *dp_cells);

/* ChainFilter::Filter (src/chain_filter.cpp:24-99).  hits is an array of
 * (query_position, database_position) pairs, filtered in place; returns the
 * number kept. */
/* BuildResults of one database chunk with one extender (the reference's
 * order: results in query order, blocks of 32 popped from the back): hits
 * are (query, hit q, hit d) triples grouped by query, queries' offsets point
 * at each query's leading delimiter (nq + 1 of them).  out: 9 ints per hit
 * (score, length, mismatches, gap openings, gap extensions, start q, start
 * d, end q, end d), in the hits' order.  Returns 0. */
int gzo_build_chunk(const uint8_t *queries, const uint64_t *query_offsets,
                    uint32_t number_queries, const uint8_t *chunk, const uint32_t *hits,
                    uint32_t number_hits, uint8_t delimiter, const int *matrix, uint32_t nl,
                    int gap_open, int gap_extension, int result_cutoff, int32_t *out);
uint32_t gzo_chain_filter(const uint8_t *query, uint32_t query_length,
                          int cutoff, const uint8_t *database,
                          uint8_t delimiter, const int *matrix,
                          uint32_t number_letters, uint32_t *hits,
                          uint32_t number_hits);

/* ReducedAlphabetVariableHashFunction::CalculateHash
 * (src/reduced_alphabet_variable_hash_function.cpp:56-86). */
int gzo_hash(const uint8_t *sequence, const uint8_t *code_map,
             uint32_t code_map_size, const int *code_scores, uint32_t max_code,
             int score_threshold, uint32_t max_length, uint32_t shift_size,
             uint32_t *hash, uint32_t *hash_length);

/* ---- the aligner -------------------------------------------------------- */

/* One database chunk: the concatenated sequence (leading delimiter, each
 * subject followed by a delimiter, database_chunk.h:412-428) plus the seed
 * index built by `db` (seed_searcher_database_parameters.cpp:41-319). */
typedef struct {
  const uint8_t *sequence;
  uint32_t sequence_length;
  uint32_t number_sequences;
  const uint32_t *offsets;        /* number_sequences + 1 */
  const char *const *names;       /* number_sequences */
  /* seed index, keyed by the sorted hash array */
  uint32_t number_hashes;
  const uint32_t *hashes;         /* ascending */
  const uint32_t *rep_offsets;    /* number_hashes + 1 */
  const uint32_t *rep_positions;  /* representative centres, per hash */
  const uint32_t *rep_member_offsets; /* number_reps + 1 */
  const uint32_t *member_positions;
  const uint32_t *value_offsets;  /* number_hashes + 1 */
  const uint32_t *value_positions;/* non-clustered positions, per hash */
  int clustering;
} gzo_db_chunk;

typedef struct {
  uint32_t number_chunks;
  const gzo_db_chunk *chunks;
  uint64_t total_length;          /* residues, Database::database_length_ */
  uint64_t total_sequences;
  uint32_t subsequence_length;    /* clustering L */
  const uint8_t *reduced_code_map;/* similarity map, 25 entries */
  /* hash function (ReducedAlphabetVariableHashFunction state) */
  const uint8_t *hash_code_map;   /* already +1 shifted, 25 entries */
  const int *hash_code_scores;    /* shifted, max_code+1 entries */
  uint32_t hash_max_code;
  int hash_score_threshold;
  uint32_t hash_max_length;
  uint32_t hash_shift_size;
  int max_mismatches;             /* GetNumberMaxMismatchs */
} gzo_database;

typedef struct {
  uint32_t number_queries;
  const uint8_t *sequences;       /* concatenation of framed query sequences */
  const uint64_t *offsets;        /* number_queries + 1 */
  const char *const *names;
  const float *lambda;            /* per-query ungapped Karlin */
  const float *K;
  int translated;                 /* 1 = TranslatedDnaQuery */
  const uint32_t *dna_lengths;    /* translated only */
} gzo_queries;

typedef struct {
  const int *matrix;
  uint32_t number_letters;
  uint8_t delimiter;
  int gap_open;
  int gap_extension;
  float normalized_ungapped_cutoff;   /* 7 */
  float normalized_gapped_trigger;    /* 22 */
  float normalized_presearch_cutoff;  /* 15 */
  float normalized_result_cutoff;     /* 25 */
  uint32_t max_results;               /* -b */
  uint32_t max_results_per_subject;   /* -v */
} gzo_parameters;

/* Aligner::Align for one query chunk (src/aligner.cpp:84-112 with
 * Presearch :170-245, BuildResults :264-348, WriteOutput): writes the
 * tabular output into a malloc'd NUL-terminated buffer (free with
 * gzo_free).  Returns 0 on success. */
int gzo_align(const gzo_queries *queries, const gzo_database *database,
              const gzo_parameters *parameters, char **output,
              size_t *output_length);

/* SeedSearcherQueryParameters::Build with one thread
 * (src/seed_searcher_query_parameters.cpp:17-219): the hash lists of the
 * queries taken as one chunk, in the order the presearch visits them
 * (src/aligner_presearch_thread.cpp:140-152); each entry's query id,
 * query-local position and check flag.  The five arrays are malloc'd
 * (gzo_free each).  Returns 0. */
int gzo_query_lists(const gzo_queries *queries, const gzo_database *database,
                    uint64_t *number_lists, uint32_t **hash, uint32_t **entry_offsets,
                    uint64_t *number_entries, uint32_t **entry_queries,
                    uint32_t **entry_positions, uint8_t **entry_checks);

/* One presearched result as AddResults sees it (PresearchedResult,
 * src/aligner_common.h:70-78). */
typedef struct {
  int chunk;
  uint32_t subject;
  int score;
  uint32_t hit_q, hit_d, start_q, start_d, end_q, end_d;
  uint32_t hit_count;
} gzo_presearched;

/* AlignerCommon::AddResults (src/aligner_common.h:147-274) for one query:
 * merges the `added` results of database chunk `chunk_id` (subject starts
 * `subject_offsets`, number_subjects + 1 of them) into `results`
 * (*number_results in, the merged count out; room for max_results).
 * `added` is sorted in place as the reference sorts it.  Returns 0. */
int gzo_add_results(const uint32_t *subject_offsets, uint32_t number_subjects, int chunk_id,
                    uint32_t max_results, uint32_t max_results_per_subject,
                    gzo_presearched *added, uint32_t number_added, gzo_presearched *results,
                    uint32_t *number_results);

/* Parity hazard 1 (src/aligner_presearch_thread.cpp:322, 363-364): 0 (the
 * default) defines the presearch end as hit + forward best offset, as the
 * GPU path does; 1 reads it as zero-initialised and accumulated over the hits
 * of one query and chunk.  Applies to later gzo_align calls on this thread. */
void gzo_set_end_reading(int reading);

/* The seed search of database chunk `chunk` alone: every query's passing
 * seeds (hits before the chain filter) in the one-thread enumeration order,
 * as (query-local position, chunk position) pairs; hits of query i are
 * pairs [query_offsets[i], query_offsets[i+1]).  Both arrays are malloc'd
 * (gzo_free).  Returns 0. */
int gzo_seed_hits(const gzo_queries *queries, const gzo_database *database,
                  const gzo_parameters *parameters, uint32_t chunk,
                  uint32_t **query_offsets, uint32_t **hits);

/* Presearch-only statistics of the last gzo_align call on this thread. */
typedef struct {
  uint64_t hash_lists;
  uint64_t distance_calculations;
  uint64_t ungapped_extensions;
  uint64_t hits;
  uint64_t hits_after_chain_filter;
  uint64_t gapped_extensions;
  /* traceback reads of op bytes not written by the same extension (their
   * value depends on earlier extensions through the same extender) */
  uint64_t stale_traceback_reads;
} gzo_stats;
void gzo_last_stats(gzo_stats *stats);

void gzo_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
