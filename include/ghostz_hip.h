/*
 * ghostz_hip.h -- C ABI of the MI355X-native GHOSTZ-GPU search path.
 *
 * Drop-in boundary for the `aln` GPU path.  Each entry point replaces one
 * reference interface, cited as file:line of akiyamalab/ghostz-gpu.  Plain C
 * types only: device pointers are `void*`/typed pointers from gz_device_malloc,
 * streams are opaque gz_stream_t handles (hipStream_t underneath), every call
 * returns a gz_status (0 = ok).  The reference exits on any CUDA error
 * (src/cuda_common.h:13-19); here the error is returned and the message is
 * available from gz_last_error() on the calling thread.
 *
 * Positions are indexes into the delimiter-framed concatenated code arrays,
 * exactly as in the reference; packed buffers are produced by
 * gz_pack_sequence_async and keep the reference's PackedAlphabetCode word
 * format (5-bit codes, 12 per uint64_t, src/packed_alphabet_code.h:15-21).
 */
#ifndef GHOSTZ_HIP_H_
#define GHOSTZ_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int gz_status;
#define GZ_OK 0
#define GZ_ERROR_INVALID_ARGUMENT 1
#define GZ_ERROR_HIP 2
#define GZ_ERROR_OUT_OF_MEMORY 3
#define GZ_ERROR_IO 4
#define GZ_ERROR_CAPACITY 5

typedef struct gz_stream_s *gz_stream_t;

/* ---- runtime (replaces src/cuda_common.h and the Thrust/pinned buffers of
 *      src/device_seeds_memory.h:18-63, src/host_seeds_memory.h:18-80) ---- */
const char *gz_last_error(void);
gz_status gz_set_device(int device);                 /* cudaSetDevice */
gz_status gz_device_malloc(void **ptr, size_t bytes);
gz_status gz_device_free(void *ptr);
gz_status gz_host_malloc(void **ptr, size_t bytes);  /* pinned */
gz_status gz_host_free(void *ptr);
gz_status gz_stream_create(gz_stream_t *stream);
gz_status gz_stream_destroy(gz_stream_t stream);
gz_status gz_stream_synchronize(gz_stream_t stream);
gz_status gz_memcpy_async(void *dst, const void *src, size_t bytes,
                          gz_stream_t stream);        /* cudaMemcpyDefault */
gz_status gz_memset_async(void *dst, int value, size_t bytes,
                          gz_stream_t stream);
/* Streaming device-to-device copy kernel (16-byte aligned): a kernel of known
 * HBM traffic, used to calibrate hardware byte counters. */
gz_status gz_copy_kernel_async(void *dst, const void *src, size_t bytes,
                               gz_stream_t stream);

/* ---- alphabet and score matrix (src/alphabet_coder.cpp:25-50,
 *      src/score_matrix_reader.cpp:57-96, src/reduced_alphabet_coder.cpp:32-98,
 *      src/align_main.cpp:102) ---- */
/* table256[byte] = protein code (A..V = 0..19, BJZ* = 20..23, else X = 24). */
void gz_protein_encode_table(uint8_t *table256);
/* BLOSUM62 as a 25 x 25 int matrix [row code][column code]. */
gz_status gz_score_matrix_blosum62(int *matrix, uint32_t *number_letters);
/* The BLOSUM62 text gz_score_matrix_blosum62 parses. */
const char *gz_score_matrix_blosum62_text(void);
/* Parse a matrix in the reference's text format into 25 x 25. */
gz_status gz_score_matrix_parse(const char *text, int *matrix,
                                uint32_t *number_letters);
/* Reduced alphabet map for the 25 protein codes (hash / similarity maps of
 * src/aligner.cpp:114-168). */
gz_status gz_reduced_code_map(const char *alphabet_sets, uint8_t *map25,
                              uint32_t *max_regular_code);

/* ---- packed alphabet (src/packed_alphabet_code.h:15-142,
 *      src/packed_alphabet_code.cu:5-48, src/aligner_gpu_data.cu:127-147) ---- */
/* Number of uint64_t words a packed buffer of `codes` codes needs. */
size_t gz_packed_words(size_t codes);
/* Pack on the device: d_codes[0..n) -> d_packed (gz_packed_words(n) words). */
gz_status gz_pack_sequence_async(const uint8_t *d_codes, size_t n,
                                 uint64_t *d_packed, gz_stream_t stream);
/* Host unpack of one code (for tests). */
uint8_t gz_packed_code_at(const uint64_t *packed, size_t position);

/* ---- DistanceCalculatorGpu (src/distance_calculator_gpu.h:18-40) ---- */
typedef struct gz_distance_calculator_s *gz_distance_calculator_t;
gz_status gz_distance_calculator_create(gz_distance_calculator_t *h);
gz_status gz_distance_calculator_destroy(gz_distance_calculator_t h);
gz_status gz_distance_calculator_set_queries(gz_distance_calculator_t h,
                                             const uint64_t *d_packed_queries);
gz_status gz_distance_calculator_set_database(gz_distance_calculator_t h,
                                              const uint64_t *d_packed_database);
gz_status gz_distance_calculator_set_reduced_code_map(
    gz_distance_calculator_t h, const uint8_t *d_reduced_code_map);
gz_status gz_distance_calculator_set_subsequence_length(
    gz_distance_calculator_t h, uint32_t length);
/* CalculateDistancesAsync (src/distance_calculator_gpu.cu:302-355): host
 * arrays may be NULL when the data is already on the device. */
gz_status gz_distance_calculator_calculate_distances_async(
    gz_distance_calculator_t h, size_t size, const uint32_t *h_query_centers,
    const uint32_t *h_database_centers, char *h_distances,
    uint32_t *d_query_centers, uint32_t *d_database_centers,
    char *d_distances, gz_stream_t stream);

/* ---- UngappedExtenderGpu (src/ungapped_extender_gpu.h:21-76) ---- */
typedef struct gz_ungapped_extender_s *gz_ungapped_extender_t;
gz_status gz_ungapped_extender_create(gz_ungapped_extender_t *h);
gz_status gz_ungapped_extender_destroy(gz_ungapped_extender_t h);
gz_status gz_ungapped_extender_set_query_seed_data_list(
    gz_ungapped_extender_t h, const uint32_t *d_query_ids,
    const uint32_t *d_query_positions);
gz_status gz_ungapped_extender_set_queries(gz_ungapped_extender_t h,
                                           uint8_t sequence_delimiter,
                                           const uint64_t *d_packed_queries,
                                           const int *d_ungapped_cutoffs,
                                           const int *d_gapped_triggers);
gz_status gz_ungapped_extender_set_database(gz_ungapped_extender_t h,
                                            const uint64_t *d_packed_database);
gz_status gz_ungapped_extender_set_score_matrix(gz_ungapped_extender_t h,
                                                const int *d_score_matrix,
                                                uint32_t number_letters);
/* Bytes of temporary device storage SetQuerySeeds needs for `size` seeds. */
size_t gz_ungapped_extender_query_seeds_temp_bytes(size_t size);
/* SetQuerySeeds (src/ungapped_extender_gpu.cu:818-852): expands
 * (query seed id, run start) pairs into per-seed query id and concatenated
 * query position with the project's own wave64 max-scan. */
gz_status gz_ungapped_extender_set_query_seeds(
    gz_ungapped_extender_t h, size_t query_seeds_size, size_t size,
    const uint32_t *h_query_seed_ids, const uint32_t *h_query_seed_starts,
    size_t temp_storage_bytes, void *d_temp_storage,
    uint32_t *d_query_seed_ids, uint32_t *d_query_seed_starts,
    uint32_t *d_query_ids, uint32_t *d_query_concatenated_positions,
    uint32_t *d_temp_array, gz_stream_t stream);
/* ExtendWithTriggerAsync (src/ungapped_extender_gpu.cu:854-899): flag[i] =
 * 1 when the ungapped X-drop score with trigger exceeds the query's trigger. */
gz_status gz_ungapped_extender_extend_with_trigger_async(
    gz_ungapped_extender_t h, size_t size, const uint32_t *h_query_ids,
    const uint32_t *h_query_concatenated_positions,
    const uint32_t *h_database_positions, char *h_flags,
    uint32_t *d_query_ids, uint32_t *d_query_concatenated_positions,
    uint32_t *d_database_positions, char *d_flags, int *d_temp_array,
    gz_stream_t stream);

/* ---- GappedExtenderGpu (src/gapped_extender_gpu.h:19-60) ---- */
typedef struct gz_gapped_extender_s *gz_gapped_extender_t;
gz_status gz_gapped_extender_create(gz_gapped_extender_t *h);
gz_status gz_gapped_extender_destroy(gz_gapped_extender_t h);
gz_status gz_gapped_extender_set_queries(gz_gapped_extender_t h,
                                         uint8_t sequence_delimiter,
                                         const uint64_t *d_packed_queries);
gz_status gz_gapped_extender_set_database(gz_gapped_extender_t h,
                                          const uint64_t *d_packed_database);
gz_status gz_gapped_extender_set_score_parameters(gz_gapped_extender_t h,
                                                  const int *d_score_matrix,
                                                  uint32_t number_letters,
                                                  int gap_open, int gap_extension,
                                                  int cutoff);
/* ConvertToGappedExtensionSeedsAsync (src/gapped_extender_gpu.cu:769-800):
 * out[i] = in[seed_positions[i]] for query and database positions. */
gz_status gz_gapped_extender_convert_to_gapped_extension_seeds_async(
    gz_gapped_extender_t h, size_t size, int reverse,
    const uint32_t *h_seed_positions, const uint32_t *h_query_positions,
    const uint32_t *h_database_positions, uint32_t *d_seed_positions,
    uint32_t *d_temp_array, uint32_t *d_query_concatenated_positions,
    uint32_t *d_database_positions, gz_stream_t stream);
/* ExtendOneSideScoreOnlyAsync (src/gapped_extender_gpu.cu:802-853) with the
 * CPU's row-major X-drop semantics (src/gapped_extender.cpp:306-471) and no
 * 128-residue limit: positions are in/out (seed in, best cell out). */
gz_status gz_gapped_extender_extend_one_side_score_only_async(
    gz_gapped_extender_t h, size_t size, int reverse,
    uint32_t *h_query_concatenated_positions, uint32_t *h_database_positions,
    int *h_scores, uint32_t *d_query_concatenated_positions,
    uint32_t *d_database_positions, int *d_scores, gz_stream_t stream);
/* Same, plus per-seed DP cell counts (the cost the learned scheduler
 * predicts); d_cells may be NULL. */
gz_status gz_gapped_extender_extend_one_side_score_only_cells_async(
    gz_gapped_extender_t h, size_t size, int reverse,
    uint32_t *d_query_concatenated_positions, uint32_t *d_database_positions,
    int *d_scores, uint32_t *d_cells, gz_stream_t stream);
/* Same, with the first wide_first seeds (predicted long) started on the
 * 64-column ring instead of the narrow one; no reference counterpart (the
 * binned batches of the learned schedule).  Results do not depend on
 * wide_first. */
/* Columns of the first LDS ring (16 or 32; default 32): bands that outgrow
 * it move to a 64-column ring, then to a global-memory ring.  No reference
 * counterpart (the reference's single ring caps sequences at 128). */
gz_status gz_gapped_extender_set_narrow_ring(gz_gapped_extender_t h, uint32_t columns);
/* Allocates the extender's workspace for calls of up to `size` jobs now
 * (calls allocate it on demand otherwise), so that a later call allocates
 * nothing and its launch sequence can be captured in a graph.  No reference
 * counterpart. */
gz_status gz_gapped_extender_reserve(gz_gapped_extender_t h, size_t size, gz_stream_t stream);
/* persistent != 0: the first ring's extensions run as one persistent launch
 * (as many waves as the device holds) fed from a device work queue: each
 * lane takes the next job as soon as its extension ends, instead of the
 * static grid where a wave waits for its slowest lane.  No reference
 * counterpart (the batched form's kernels); results do not depend on it. */
gz_status gz_gapped_extender_set_persistent(gz_gapped_extender_t h, int persistent);
/* Seeds of the last run that outgrew the narrow ring / the 64-column ring. */
gz_status gz_gapped_extender_overflows(gz_gapped_extender_t h, uint32_t *narrow,
                                       uint32_t *ring64);
gz_status gz_gapped_extender_extend_one_side_score_only_binned_async(
    gz_gapped_extender_t h, size_t size, size_t wide_first, int reverse,
    uint32_t *d_query_concatenated_positions, uint32_t *d_database_positions,
    int *d_scores, uint32_t *d_cells, gz_stream_t stream);
/* Same, with the jobs run in the order d_order gives (slot k runs job
 * d_order[k], a permutation of [0, size)); each job reads and writes its own
 * index, so a caller that orders jobs needs no gathered copies.  No
 * reference counterpart (the learned schedule's order).  An entry of
 * d_order at or past size is skipped and counted as a failure. */
gz_status gz_gapped_extender_extend_one_side_score_only_ordered_async(
    gz_gapped_extender_t h, size_t size, int reverse, const uint32_t *d_order,
    uint32_t *d_query_concatenated_positions, uint32_t *d_database_positions,
    int *d_scores, uint32_t *d_cells, gz_stream_t stream);
/* The same as gz_gapped_extender_extend_one_side_score_only_cells_async
 * (d_cells may be NULL), allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops).  Its workspace is one static device
 * array per GPU (about 18 MB, held until the process ends); the call
 * rejects more than 1,048,576 jobs or a maximum sequence length
 * (gz_gapped_extender_set_max_sequence_length) above 4094 residues with
 * GZ_ERROR_CAPACITY.  Every call on a GPU uses that GPU's one array, so
 * calls that overlap on one GPU (two streams, two threads) race: serialize
 * them.  gz_gapped_extender_failures reports the last call's failures. */
gz_status gz_gapped_extender_extend_one_side_score_only_noalloc_async(
    gz_gapped_extender_t h, size_t size, int reverse,
    uint32_t *d_query_concatenated_positions, uint32_t *d_database_positions,
    int *d_scores, uint32_t *d_cells, gz_stream_t stream);
/* Longest query the extender must handle (sizes the wide-band fallback
 * ring; default 65534 residues).  No counterpart in the reference, whose
 * GPU kernel caps the query side at 128 residues
 * (src/gapped_extender_gpu.h:25). */
gz_status gz_gapped_extender_set_max_sequence_length(gz_gapped_extender_t h,
                                                     uint32_t length);
/* Extensions of the last call that no ring could hold (must be 0; a
 * non-zero count means those results are invalid).  Synchronizes. */
gz_status gz_gapped_extender_failures(gz_gapped_extender_t h,
                                      uint32_t *failures);

/* ---- HostSeedsMemory (src/host_seeds_memory.h:18-80,
 *      src/host_seeds_memory.cpp:26-37): six pinned arrays of `size`
 *      entries that a GpuStreamController run reads and writes ---- */
typedef struct gz_host_seeds_memory_s *gz_host_seeds_memory_t;
gz_status gz_host_seeds_memory_create(size_t size, gz_host_seeds_memory_t *out);
gz_status gz_host_seeds_memory_destroy(gz_host_seeds_memory_t m);
size_t gz_host_seeds_memory_size(gz_host_seeds_memory_t m);
uint32_t *gz_host_seeds_memory_query_seed_ids(gz_host_seeds_memory_t m);
uint32_t *gz_host_seeds_memory_query_seed_starts(gz_host_seeds_memory_t m);
uint32_t *gz_host_seeds_memory_query_ids(gz_host_seeds_memory_t m);
uint32_t *gz_host_seeds_memory_query_concatenated_positions(gz_host_seeds_memory_t m);
uint32_t *gz_host_seeds_memory_database_positions(gz_host_seeds_memory_t m);
int *gz_host_seeds_memory_result_values(gz_host_seeds_memory_t m);

/* ---- GpuStreamController (src/gpu_stream_controller.h:24-101,
 *      src/gpu_stream_controller.cpp:20-77) with GpuStreamRunner
 *      (src/gpu_stream_runner.cpp:14-107): two resources, each two streams
 *      and a DeviceSeedsMemory (src/device_seeds_memory.h:18-63).  A Run*
 *      call takes a free resource (GZ_ERROR_INVALID_ARGUMENT when none is
 *      free, where the reference asserts), enqueues copies and kernels on
 *      it, and returns its id; results are in the host memory after
 *      WaitRun(id); ReleaseResource(id) hands the resource back.  Gapped
 *      runs are serialized on the device (the extender's ring workspace is
 *      per handle); distance and ungapped runs of the two resources overlap.
 *      Result values: distance runs write one char per seed, ungapped runs one
 *      flag char per seed, gapped runs an int score per seed with the
 *      positions replaced by the best cells. ---- */
typedef struct gz_stream_controller_s *gz_stream_controller_t;
gz_status gz_stream_controller_create(gz_stream_controller_t *out);
gz_status gz_stream_controller_destroy(gz_stream_controller_t c);
/* Init(device_buffer_size): seed arrays of that many entries per resource. */
gz_status gz_stream_controller_init(gz_stream_controller_t c, size_t device_buffer_size);
gz_status gz_stream_controller_set_distance_calculator(gz_stream_controller_t c,
                                                       gz_distance_calculator_t d);
gz_status gz_stream_controller_set_ungapped_extender(gz_stream_controller_t c,
                                                     gz_ungapped_extender_t u);
gz_status gz_stream_controller_set_gapped_extender(gz_stream_controller_t c,
                                                   gz_gapped_extender_t g);
/* Run(DistanceCalculationSeedsParameters{size}, host memory): distances of
 * (query_concatenated_positions[i], database_positions[i]) into
 * ((char *)result_values)[i]. */
gz_status gz_stream_controller_run_distance_calculation(gz_stream_controller_t c, size_t size,
                                                        gz_host_seeds_memory_t m,
                                                        int *resource_id);
/* Run(UngappedExtensionWithTriggerSeedsParameters{query_seed_size, size}):
 * query_seed_ids / query_seed_starts hold the seed runs, database_positions
 * the size seeds; ((char *)result_values)[i] = trigger flag. */
gz_status gz_stream_controller_run_ungapped_extension_with_trigger(gz_stream_controller_t c,
                                                                   size_t query_seed_size,
                                                                   size_t size,
                                                                   gz_host_seeds_memory_t m,
                                                                   int *resource_id);
/* Run(GappedExtensionSeesParameters): per side, query_seed_ids[off + i]
 * indexes the side's positions (ConvertToGappedExtensionSeedsAsync), then
 * ExtendOneSideScoreOnlyAsync writes result_values[off + i] and the best
 * cells into the position arrays. */
gz_status gz_stream_controller_run_gapped_extension(gz_stream_controller_t c,
                                                    size_t reverse_offset, size_t reverse_length,
                                                    size_t forward_offset, size_t forward_length,
                                                    gz_host_seeds_memory_t m, int *resource_id);
/* WaitRun(id) (id < 0: nothing to wait for) and WaitRun() for all. */
gz_status gz_stream_controller_wait_run(gz_stream_controller_t c, int resource_id);
gz_status gz_stream_controller_wait_run_all(gz_stream_controller_t c);
/* FinishedRun(id): 1 when the resource's streams are idle (or id < 0). */
int gz_stream_controller_finished_run(gz_stream_controller_t c, int resource_id);
gz_status gz_stream_controller_release_resource(gz_stream_controller_t c, int resource_id);
uint32_t gz_stream_controller_available_resources(gz_stream_controller_t c);

/* ---- queries (Queries / ProteinQuery / TranslatedDnaQuery,
 *      src/queries.cpp:183-259, src/protein_query.cpp:22-60,
 *      src/translated_dna_query.cpp:27-85) ---- */
typedef struct gz_queries_s *gz_queries_t;
typedef struct {
  int translated; /* -q d -t p */
  int filter;     /* -F T: SEG */
  int threads;    /* 0 = all cores */
} gz_query_options;
typedef struct {
  uint64_t number_queries;
  const uint8_t *codes;    /* framed queries back to back */
  const uint64_t *offsets; /* number_queries + 1 */
  const char *const *names;
  const float *lambda;     /* per-query ungapped Karlin */
  const float *K;
  int translated;
  const uint32_t *dna_lengths;
} gz_queries_view_t;
gz_status gz_queries_create(const char *const *names, const char *residues,
                            const uint64_t *offsets, uint64_t n,
                            const gz_query_options *options, gz_queries_t *out);
gz_status gz_queries_read_fasta(const char *path, const gz_query_options *options,
                                gz_queries_t *out);
gz_status gz_queries_destroy(gz_queries_t q);
gz_status gz_queries_view(gz_queries_t q, gz_queries_view_t *view);


/* ---- database (Database / DatabaseChunk / SeedSearcherDatabaseParameters,
 *      src/database.h:440-514, src/database_chunk.h:188-210,
 *      src/seed_searcher_database_parameters.cpp:41-319) ---- */
typedef struct gz_database_s *gz_database_t;
typedef struct {
  int clustering;              /* -C T */
  uint32_t subsequence_length; /* -L, 0 = 10 */
  int seed_threshold;          /* -s, 0 = 39 */
  uint64_t chunk_size;         /* -l residues, 0 = 1 GiB */
  int threads;
} gz_database_options;
typedef struct {
  uint32_t number_chunks;
  uint64_t total_length;
  uint64_t total_sequences;
  uint32_t max_length;
  uint32_t subsequence_length;
  uint8_t reduced_code_map[25];
  uint8_t hash_code_map[25];   /* shifted by one */
  int hash_code_scores[16];
  uint32_t hash_max_code;
  int hash_score_threshold;
  uint32_t hash_max_length;
  uint32_t hash_shift;
  int max_mismatches;
} gz_database_info_t;
typedef struct {
  const uint8_t *sequence;
  uint32_t sequence_length;
  uint32_t number_sequences;
  const uint32_t *offsets;
  const char *const *names;
  uint32_t number_hashes;
  const uint32_t *hashes, *rep_offsets, *rep_positions, *rep_member_offsets,
      *member_positions, *value_offsets, *value_positions;
  uint64_t number_reps, number_members, number_values;
  int clustering;
} gz_database_chunk_view_t;
gz_status gz_database_create(const char *const *names, const char *residues,
                             const uint64_t *offsets, uint64_t n,
                             const gz_database_options *options, gz_database_t *out);
gz_status gz_database_read_fasta(const char *path, const gz_database_options *options,
                                 gz_database_t *out);
/* Database::Load (src/database.h:229-237, :517-538) with every chunk
 * resident (DatabaseChunk::Load, src/database_chunk.h:303-400): reads
 * <prefix>.inf and <prefix>_<i>.{inf,off,nam,seq,scp,sdp}; GZ_ERROR_IO with
 * the file named when one is missing, truncated or inconsistent. */
gz_status gz_database_load(const char *prefix, gz_database_t *out);
/* Database::Save (src/database.h:240-251, :540-554;
 * DatabaseChunk::Save, src/database_chunk.h:267-278): the same file set. */
gz_status gz_database_save(gz_database_t db, const char *prefix);
gz_status gz_database_destroy(gz_database_t db);
gz_status gz_database_info(gz_database_t db, gz_database_info_t *info);
gz_status gz_database_chunk(gz_database_t db, uint32_t chunk,
                            gz_database_chunk_view_t *view);

/* Query hash lists of a query set taken as one query chunk, as
 * SeedSearcherGpuQueryParameters::Build forms them
 * (src/seed_searcher_gpu_query_parameters.cpp:17-272 over
 * SeedSearcherQueryParameters::Build, src/seed_searcher_query_parameters.cpp
 * :17-219): lists in the presearch's visiting order, entries with their
 * query, position in the shared-delimiter concatenation of the queries and
 * the distance-check flag.  Built on the current device (on_device != 0) or
 * on the host; the arrays are malloc'd (gz_query_lists_free). */
typedef struct {
  uint64_t number_lists;
  uint64_t number_entries;
  uint32_t *hash;          /* number_lists */
  uint32_t *entry_offsets; /* number_lists + 1 */
  uint32_t *entry_positions;
  uint32_t *entry_queries;
  uint8_t *entry_checks;
  double seconds;          /* build time (device: HIP-event time of the kernels) */
} gz_query_lists;
gz_status gz_query_lists_build(gz_queries_t q, gz_database_t db, int on_device,
                               gz_query_lists *out);
void gz_query_lists_free(gz_query_lists *lists);

/* ---- aligner (AlignerGpu, src/aligner_gpu.h:22-62; one device, one
 *      GpuStreamController-style pipeline) ---- */
typedef struct gz_aligner_s *gz_aligner_t;
#define GZ_SCHEDULE_PLAIN 0         /* gapped extensions in hit order */
#define GZ_SCHEDULE_LENGTH_SORTED 1 /* by distance to the delimiter */
#define GZ_SCHEDULE_LEARNED 2       /* by the learned cost model */
typedef struct {
  int gap_open, gap_extension;               /* 11 / 1 */
  float normalized_ungapped_cutoff;          /* 7 bits */
  float normalized_gapped_trigger;           /* 22 bits */
  float normalized_presearch_cutoff;         /* 15 bits */
  float normalized_result_cutoff;            /* 25 bits */
  uint32_t max_results;                      /* -b */
  uint32_t max_results_per_subject;          /* -v */
  uint64_t queries_chunk_size;               /* -l */
  int threads;                               /* host threads, 0 = all */
  int schedule;                              /* GZ_SCHEDULE_* */
  int record_traces;                         /* keep per-extension DP cells */
  int batched;              /* pipeline query blocks: host stages of one block
                               overlap the device stages of the next */
  uint32_t pipeline_blocks; /* query blocks per chunk when batched (0 = 8) */
} gz_aligner_parameters;
typedef struct {
  double seconds_total, seconds_query_lists, seconds_upload, seconds_seed,
      seconds_chain, seconds_gapped, seconds_add, seconds_build, seconds_output;
  uint64_t query_chunks, lists, units_rep, units_value, seed_hits, chained_hits,
      gapped_extensions, gapped_cells, results, sequential_build_chunks;
  double gapped_kernel_ms;     /* HIP events around the gapped launches */
  double seed_rep_kernel_ms;   /* HIP events around the seed kernels */
  double seed_value_kernel_ms;
  uint64_t seed_rep_launches, seed_value_launches, gapped_launches;
  double schedule_kernel_ms;   /* device ordering of gapped jobs (features,
                                  cost model, counting sort, gathers) */
  double traceback_kernel_ms;  /* device result build */
  uint64_t traceback_host_jobs; /* results the device left to the host */
  uint64_t gated_jobs;          /* learned schedule: jobs the cost model
                                   predicted (query side >= the gate) */
} gz_align_stats;
typedef struct {
  uint32_t query_position; /* concatenated seed position */
  uint32_t database_position;
  uint32_t chunk;
  uint32_t reverse;
  uint32_t cells;          /* DP cells the extension evaluated */
} gz_gapped_trace;
void gz_aligner_default_parameters(gz_aligner_parameters *p);
/* Per-query ungapped cutoffs and gapped triggers, and the gapped
 * extension's cutoff, as AlignerGpu::SetQueriesData computes them
 * (src/aligner_gpu.cpp:213-242, src/statistics.cpp:28-49,236-249) from the
 * normalized values in p (NULL: defaults). */
gz_status gz_queries_cutoffs(gz_queries_t q, const gz_aligner_parameters *p,
                             int *ungapped_cutoffs, int *gapped_triggers,
                             int *gapped_cutoff);
gz_status gz_aligner_create(const gz_aligner_parameters *p, gz_aligner_t *out);
gz_status gz_aligner_destroy(gz_aligner_t a);
gz_status gz_aligner_set_database(gz_aligner_t a, gz_database_t db);
/* The batched form launches each query-length bucket's gapped sequence as a
 * HIP graph (captured once per bucket and configuration, then updated with
 * each call's arguments): how many graphs were instantiated and how many
 * launches reused one through an update.  No reference counterpart. */
gz_status gz_aligner_graph_counts(gz_aligner_t a, uint64_t *instantiations, uint64_t *updates);
/* The batched form's autotuner (no reference counterpart): per query-length
 * bucket b (<= 128, <= 512, <= 2048, longer residues), chosen[b] = the
 * gapped configuration it uses next (-1: bucket not seen; configuration c:
 * first ring 16 columns when c & 1, else 32; static grid when c & 2, else
 * the persistent work queue), and for each configuration the measured
 * kernel milliseconds per job and trials (4 x 4 arrays, bucket-major).  Any
 * pointer may be NULL. */
gz_status gz_aligner_tuner_state(gz_aligner_t a, int *chosen, double *ms_per_job,
                                 uint32_t *trials);
/* Flat fp32 weights of the cost model (embed, w1, b1, w2, b2, w3, b3). */
gz_status gz_aligner_set_cost_model(gz_aligner_t a, const float *weights, size_t count);
/* AlignerGpu::Align for in-memory queries: presearch, BuildResults and
 * WriteOutput; *output (NUL-terminated, *output_length bytes before the NUL)
 * belongs to the caller until gz_free. */
gz_status gz_aligner_align(gz_aligner_t a, gz_queries_t q, char **output,
                           size_t *output_length, gz_align_stats *stats);
uint64_t gz_aligner_traces(gz_aligner_t a, gz_gapped_trace *out, uint64_t capacity);
gz_status gz_aligner_last_concatenated_queries(gz_aligner_t a, const uint8_t **codes,
                                               uint64_t *length);
/* `ghostz-gpu aln -i <queries> -d <database prefix> -o <output>`
 * (AlignMain::Run, src/align_main.cpp:34-80): the database files of
 * gz_database_save / `ghostz-gpu db`; a missing query file or <prefix>.inf
 * is GZ_ERROR_IO with the reference's message. */
gz_status gz_align_main(const char *queries_path, const char *database_prefix,
                        const char *output_path, int translated,
                        const gz_aligner_parameters *p);
/* `ghostz-gpu aln` on FASTA inputs (the database is built in memory). */
gz_status gz_align_fasta(const char *queries_path, const char *database_path,
                         const char *output_path, int translated,
                         const gz_aligner_parameters *p,
                         const gz_database_options *database_options);
void gz_free(void *p);

/* ---- SeedSearcherGpu (src/seed_searcher_gpu.h:28-115,
 *      src/seed_searcher_gpu.cpp:48-548) on the device seed stage.
 *      The reference builds host seed lists that its kernels score (distance,
 *      ungapped extension with trigger) before the host sorts the passing
 *      seeds into per-query buckets (src/aligner_gpu_presearch_thread.cpp
 *      :822-896).  Here the searches run that whole chain on the device for
 *      the queries [query_begin, query_end) (query_begin a multiple of 64,
 *      query_end one too or the number of queries) and return its output:
 *      the passing seeds of each query in the one-thread CPU enumeration
 *      order (lists in presearch order; per list the representatives, each
 *      followed by its members, then the non-clustered values), which is the
 *      order the chain filter expects.  Calls on one searcher are serialized
 *      by the caller. ---- */
typedef struct gz_seed_searcher_s *gz_seed_searcher_t;
typedef struct {
  uint64_t number_hits;
  uint32_t number_queries;              /* query_end - query_begin */
  const uint32_t *query_offsets;        /* number_queries + 1: hits of query
                                           query_begin + i are [off[i], off[i+1]) */
  const uint32_t *query_positions;      /* query-local (Hit::query_sequence_position) */
  const uint32_t *database_positions;   /* chunk positions */
} gz_seed_hits;  /* owned by the searcher, valid until its next call */
/* p: the cutoffs (NULL: defaults); the searcher owns its device state. */
gz_status gz_seed_searcher_create(const gz_aligner_parameters *p, gz_seed_searcher_t *out);
gz_status gz_seed_searcher_destroy(gz_seed_searcher_t h);
/* Reset(QueryParameters&, DatabaseParameters&): the query set (one query
 * chunk) and database chunk `chunk` become the searched pair; the hash lists
 * are built on the device (again only when the query set changes).  Returns
 * the number of hash lists. */
gz_status gz_seed_searcher_reset(gz_seed_searcher_t h, gz_queries_t queries,
                                 gz_database_t database, uint32_t chunk,
                                 uint32_t *number_hash_lists);
/* SearchSeedsFromRepresentativesForSimilarityFiltering(id, seed list): the
 * distance-calculation pairs of hash list `list` (presearch order):
 * (checked entry's concatenated query position, representative position),
 * representatives outer.  *count receives the number of pairs;
 * GZ_ERROR_CAPACITY when it exceeds `capacity`.  Unclustered chunks have
 * none. */
gz_status gz_seed_searcher_search_seeds_from_representatives_for_similarity_filtering(
    gz_seed_searcher_t h, uint32_t list, uint32_t *query_concatenated_positions,
    uint32_t *database_positions, size_t capacity, size_t *count);
/* SearchSeedsFromClusteringSubsequence: passing seeds from the
 * representatives and their members (distance filter included). */
gz_status gz_seed_searcher_search_seeds_from_clustering_subsequence(gz_seed_searcher_t h,
                                                                    uint32_t query_begin,
                                                                    uint32_t query_end,
                                                                    gz_seed_hits *out);
/* SearchSeedsFromNonClusteringSubsequence: passing seeds from the
 * non-clustered values. */
gz_status gz_seed_searcher_search_seeds_from_non_clustering_subsequence(gz_seed_searcher_t h,
                                                                        uint32_t query_begin,
                                                                        uint32_t query_end,
                                                                        gz_seed_hits *out);
/* Both sources, merged in the enumeration order. */
gz_status gz_seed_searcher_search_seeds(gz_seed_searcher_t h, uint32_t query_begin,
                                        uint32_t query_end, gz_seed_hits *out);

/* ---- result build (AlignerBuildResultsThread, aligner_build_results_thread.h
 *      :64-190 with GappedExtender's traceback, gapped_extender.cpp:70-304,473-531,
 *      and AlignerCommon::BuildResult, aligner_common.h:276-343) ---- */
typedef struct {
  uint32_t query_offset; /* concatenated position of the query's leading delimiter */
  uint32_t query_length; /* framed length */
  uint32_t hit_q, hit_d; /* query-local / chunk positions of the hit */
} gz_traceback_job;
#define GZ_TRACEBACK_STALE 1    /* read a byte the extension did not write */
#define GZ_TRACEBACK_OVERFLOW 2 /* beyond the device scratch: build on the host */
typedef struct {
  int score;
  uint32_t start_q, start_d, end_q, end_d;
  uint32_t length, mismatches, gap_openings, gap_extensions;
  uint32_t flags;
} gz_traceback_result;
/* Re-extends both sides of every job with traceback at `cutoff` on the
 * device and counts the alignment; synchronous. */
gz_status gz_build_results(const uint64_t *d_packed_queries, const uint64_t *d_packed_database,
                           const gz_traceback_job *jobs, size_t n, uint32_t max_query_length,
                           uint32_t max_subject_length, const int *d_matrix,
                           uint32_t number_letters, int gap_open, int gap_extension, int cutoff,
                           gz_traceback_result *results, gz_stream_t stream);
/* The same for one job on the host (BLOSUM62 11/1), from unpacked codes. */
gz_status gz_build_result_host(const uint8_t *query, uint32_t query_length,
                               const uint8_t *database, uint32_t hit_q, uint32_t hit_d,
                               int cutoff, gz_traceback_result *result);

/* ---- learned extension-cost model (no counterpart in the reference: it
 *      only orders gapped extensions for load balance) ----
 * Flat fp32 layout: embed[21][8], w1[256][1024], b1[256], w2[64][256],
 * b2[64], w3[64], b3[1].  Inputs are 128 residue codes per seed in [0, 21). */
typedef struct gz_mlp_s *gz_mlp_t;
size_t gz_mlp_weight_count(void);
gz_status gz_mlp_init_weights(uint64_t seed, float *flat);
gz_status gz_mlp_create(const float *flat, size_t count, gz_mlp_t *out);
gz_status gz_mlp_destroy(gz_mlp_t m);
gz_status gz_mlp_get_weights(gz_mlp_t m, float *flat);
/* Predicted log2 DP cells; kernel_ms (optional) = device forward time. */
gz_status gz_mlp_predict(gz_mlp_t m, const uint8_t *codes, size_t n, float *out,
                         float *kernel_ms);
/* The same for each row with its two halves of windows swapped (the other
 * side of the same hit), without moving the codes: the aligner predicts a
 * hit's forward side from its reverse side's features this way.  No
 * reference counterpart. */
gz_status gz_mlp_predict_swapped(gz_mlp_t m, const uint8_t *codes, size_t n, float *out);
/* One MSE step: gradients (flat layout, optional out) and Adam (lr > 0). */
gz_status gz_mlp_train_step(gz_mlp_t m, const uint8_t *codes, const float *targets,
                            size_t n, float lr, float *loss, float *flat_grad);
/* The same steps over a training set held on the device: codes (n x 128)
 * and targets uploaded once; each whole batch of `order` (row indices, for
 * example several epochs' permutations one after the other) is one step
 * with Adam at lr, with no host round trip between steps; *mean_loss = mean
 * of the steps' losses, *kernel_ms = device time of the steps. */
gz_status gz_mlp_train_epochs(gz_mlp_t m, const uint8_t *codes, const float *targets, size_t n,
                              const uint32_t *order, size_t order_length, size_t batch, float lr,
                              float *mean_loss, float *kernel_ms);
/* 128 feature codes per seed from unpacked concatenated query / chunk codes. */
void gz_cost_features(const uint8_t *queries, const uint8_t *database,
                      const uint32_t *query_positions, const uint32_t *database_positions,
                      size_t n, int reverse, uint8_t *out);
/* The same codes computed on the device from packed sequences (all pointers
 * are device pointers; asynchronous on `stream`). */
gz_status gz_cost_features_packed_async(const uint64_t *d_packed_queries,
                                        const uint64_t *d_packed_database,
                                        const uint32_t *d_query_positions,
                                        const uint32_t *d_database_positions, size_t n,
                                        int reverse, uint8_t *d_out, gz_stream_t stream);

/* ---- host preparation pieces (tests) ---- */
int gz_seg_mask(const char *sequence, char *masked);    /* SequenceSegFilter */
double gz_seg_ln_factorial(int n);
int gz_translate(const char *dna, char *frames, uint32_t *lengths); /* Translator */
void gz_karlin_block(const double *p_at_min, int min, int max, float *lambda,
                     float *K, float *H);                /* BlastKarlinBlkCalc */
int gz_query_karlin(const uint8_t *framed, uint32_t length, float *lambda, float *K,
                    float *H);
uint64_t gz_search_space(uint64_t query_length, uint64_t db_length,
                         uint64_t db_sequences);

#ifdef __cplusplus
}
#endif
#endif /* GHOSTZ_HIP_H_ */
