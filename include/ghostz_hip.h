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

#ifdef __cplusplus
}
#endif
#endif /* GHOSTZ_HIP_H_ */
