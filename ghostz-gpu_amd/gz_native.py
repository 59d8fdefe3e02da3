"""ctypes binding of libghostz_amd.so (the C ABI declared in include/ghostz_hip.h).

The library is the product: HIP kernels for gfx950 plus the C++ host side.
Loading fails loudly when it has not been built -- there is no fallback.
"""
import ctypes
import os
import time
from ctypes import POINTER, c_char_p, c_int, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libghostz_amd.so")

_lib = None

u8p = POINTER(c_uint8)
u32p = POINTER(c_uint32)
u64p = POINTER(c_uint64)
intp = POINTER(c_int)

# name -> (restype, argtypes)
_SIGNATURES = {
    "gz_last_error": (c_char_p, []),
    "gz_set_device": (c_int, [c_int]),
    "gz_device_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "gz_device_free": (c_int, [c_void_p]),
    "gz_host_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "gz_host_free": (c_int, [c_void_p]),
    "gz_stream_create": (c_int, [POINTER(c_void_p)]),
    "gz_stream_destroy": (c_int, [c_void_p]),
    "gz_stream_synchronize": (c_int, [c_void_p]),
    "gz_memcpy_async": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "gz_memset_async": (c_int, [c_void_p, c_int, c_size_t, c_void_p]),
    "gz_copy_kernel_async": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "gz_protein_encode_table": (None, [c_void_p]),
    "gz_score_matrix_blosum62": (c_int, [c_void_p, u32p]),
    "gz_score_matrix_blosum62_text": (c_char_p, []),
    "gz_score_matrix_parse": (c_int, [c_char_p, c_void_p, u32p]),
    "gz_reduced_code_map": (c_int, [c_char_p, c_void_p, u32p]),
    "gz_packed_words": (c_size_t, [c_size_t]),
    "gz_pack_sequence_async": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p]),
    "gz_packed_code_at": (c_uint8, [c_void_p, c_size_t]),
    "gz_distance_calculator_create": (c_int, [POINTER(c_void_p)]),
    "gz_distance_calculator_destroy": (c_int, [c_void_p]),
    "gz_distance_calculator_set_queries": (c_int, [c_void_p, c_void_p]),
    "gz_distance_calculator_set_database": (c_int, [c_void_p, c_void_p]),
    "gz_distance_calculator_set_reduced_code_map": (c_int, [c_void_p, c_void_p]),
    "gz_distance_calculator_set_subsequence_length": (c_int, [c_void_p, c_uint32]),
    "gz_distance_calculator_calculate_distances_async": (
        c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p]),
    "gz_ungapped_extender_create": (c_int, [POINTER(c_void_p)]),
    "gz_ungapped_extender_destroy": (c_int, [c_void_p]),
    "gz_ungapped_extender_set_query_seed_data_list": (c_int, [c_void_p, c_void_p, c_void_p]),
    "gz_ungapped_extender_set_queries": (c_int, [c_void_p, c_uint8, c_void_p, c_void_p,
                                                 c_void_p]),
    "gz_ungapped_extender_set_database": (c_int, [c_void_p, c_void_p]),
    "gz_ungapped_extender_set_score_matrix": (c_int, [c_void_p, c_void_p, c_uint32]),
    "gz_ungapped_extender_query_seeds_temp_bytes": (c_size_t, [c_size_t]),
    "gz_ungapped_extender_set_query_seeds": (
        c_int, [c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p,
                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gz_ungapped_extender_extend_with_trigger_async": (
        c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gz_gapped_extender_create": (c_int, [POINTER(c_void_p)]),
    "gz_gapped_extender_destroy": (c_int, [c_void_p]),
    "gz_gapped_extender_set_queries": (c_int, [c_void_p, c_uint8, c_void_p]),
    "gz_gapped_extender_set_database": (c_int, [c_void_p, c_void_p]),
    "gz_gapped_extender_set_score_parameters": (c_int, [c_void_p, c_void_p, c_uint32, c_int,
                                                        c_int, c_int]),
    "gz_gapped_extender_convert_to_gapped_extension_seeds_async": (
        c_int, [c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p, c_void_p]),
    "gz_gapped_extender_extend_one_side_score_only_async": (
        c_int, [c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p]),
    "gz_gapped_extender_extend_one_side_score_only_cells_async": (
        c_int, [c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gz_gapped_extender_extend_one_side_score_only_binned_async": (
        c_int, [c_void_p, c_size_t, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p]),
    "gz_gapped_extender_extend_one_side_score_only_ordered_async": (
        c_int, [c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p]),
    "gz_gapped_extender_extend_one_side_score_only_noalloc_async": (
        c_int, [c_void_p, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "gz_gapped_extender_set_max_sequence_length": (c_int, [c_void_p, c_uint32]),
    "gz_gapped_extender_set_narrow_ring": (c_int, [c_void_p, c_uint32]),
    "gz_gapped_extender_set_persistent": (c_int, [c_void_p, c_int]),
    "gz_gapped_extender_reserve": (c_int, [c_void_p, c_size_t, c_void_p]),
    "gz_gapped_extender_overflows": (c_int, [c_void_p, u32p, u32p]),
    "gz_gapped_extender_failures": (c_int, [c_void_p, u32p]),
}


class GzError(RuntimeError):
    pass


def lib():
    """Load libghostz_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GzError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status):
    if status != 0:
        msg = lib().gz_last_error()
        raise GzError(f"gz status {status}: {msg.decode() if msg else ''}")


def ptr(a):
    """Address of a contiguous numpy array (host) as c_void_p."""
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---- host helpers -----------------------------------------------------------
def encode_table():
    t = np.zeros(256, np.uint8)
    lib().gz_protein_encode_table(ptr(t))
    return t


def encode_protein(s):
    return encode_table()[np.frombuffer(s.encode() if isinstance(s, str) else s, np.uint8)]


def blosum62():
    m = np.zeros(25 * 25, np.int32)
    nl = c_uint32()
    check(lib().gz_score_matrix_blosum62(ptr(m), ctypes.byref(nl)))
    return m, nl.value


def blosum62_text():
    return lib().gz_score_matrix_blosum62_text()


def reduced_map(sets="A KR EDNQ C G H ILVM FYW P ST"):
    m = np.zeros(25, np.uint8)
    mx = c_uint32()
    check(lib().gz_reduced_code_map(sets.encode(), ptr(m), ctypes.byref(mx)))
    return m, mx.value


# ---- device helpers ---------------------------------------------------------
class DeviceBuffer:
    """Owning device allocation (gz_device_malloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = c_void_p()
        check(lib().gz_device_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_array(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().gz_memcpy_async(b.ptr, ptr(a), a.nbytes, stream))
            check(lib().gz_stream_synchronize(stream))
        return b

    def to_array(self, dtype, count, stream=None):
        out = np.empty(count, dtype)
        if out.nbytes:
            check(lib().gz_memcpy_async(ptr(out), self.ptr, out.nbytes, stream))
            check(lib().gz_stream_synchronize(stream))
        return out

    def free(self):
        if self.ptr:
            lib().gz_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = c_void_p()
        check(lib().gz_stream_create(ctypes.byref(p)))
        self.handle = p

    def sync(self):
        check(lib().gz_stream_synchronize(self.handle))

    def __del__(self):
        try:
            lib().gz_stream_destroy(self.handle)
        except Exception:
            pass


def pack_on_device(codes, stream=None):
    """Upload uint8 codes and pack them on the device; returns (codes_buf, packed_buf)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    d_codes = DeviceBuffer.from_array(codes, stream)
    words = lib().gz_packed_words(codes.size)
    d_packed = DeviceBuffer(words * 8)
    check(lib().gz_pack_sequence_async(d_codes.ptr, codes.size, d_packed.ptr, stream))
    check(lib().gz_stream_synchronize(stream))
    return d_codes, d_packed


# ---- pipeline (queries, database, aligner) ----------------------------------
from ctypes import Structure, c_char, c_double, c_float  # noqa: E402


class QueryOptions(Structure):
    _fields_ = [("translated", c_int), ("filter", c_int), ("threads", c_int)]


class QueriesView(Structure):
    _fields_ = [("number_queries", c_uint64), ("codes", c_void_p), ("offsets", c_void_p),
                ("names", c_void_p), ("lambda_", c_void_p), ("K", c_void_p),
                ("translated", c_int), ("dna_lengths", c_void_p)]


class DatabaseOptions(Structure):
    _fields_ = [("clustering", c_int), ("subsequence_length", c_uint32),
                ("seed_threshold", c_int), ("chunk_size", c_uint64), ("threads", c_int)]


class DatabaseInfo(Structure):
    _fields_ = [("number_chunks", c_uint32), ("total_length", c_uint64),
                ("total_sequences", c_uint64), ("max_length", c_uint32),
                ("subsequence_length", c_uint32), ("reduced_code_map", c_uint8 * 25),
                ("hash_code_map", c_uint8 * 25), ("hash_code_scores", c_int * 16),
                ("hash_max_code", c_uint32), ("hash_score_threshold", c_int),
                ("hash_max_length", c_uint32), ("hash_shift", c_uint32),
                ("max_mismatches", c_int)]


class ChunkView(Structure):
    _fields_ = [("sequence", c_void_p), ("sequence_length", c_uint32),
                ("number_sequences", c_uint32), ("offsets", c_void_p), ("names", c_void_p),
                ("number_hashes", c_uint32), ("hashes", c_void_p), ("rep_offsets", c_void_p),
                ("rep_positions", c_void_p), ("rep_member_offsets", c_void_p),
                ("member_positions", c_void_p), ("value_offsets", c_void_p),
                ("value_positions", c_void_p), ("number_reps", c_uint64),
                ("number_members", c_uint64), ("number_values", c_uint64),
                ("clustering", c_int)]


class AlignerParameters(Structure):
    _fields_ = [("gap_open", c_int), ("gap_extension", c_int),
                ("normalized_ungapped_cutoff", c_float), ("normalized_gapped_trigger", c_float),
                ("normalized_presearch_cutoff", c_float), ("normalized_result_cutoff", c_float),
                ("max_results", c_uint32), ("max_results_per_subject", c_uint32),
                ("queries_chunk_size", c_uint64), ("threads", c_int), ("schedule", c_int),
                ("record_traces", c_int), ("batched", c_int), ("pipeline_blocks", c_uint32)]


class AlignStats(Structure):
    _fields_ = [(n, c_double) for n in (
        "seconds_total", "seconds_query_lists", "seconds_upload", "seconds_seed",
        "seconds_chain", "seconds_gapped", "seconds_add", "seconds_build",
        "seconds_output")] + [(n, c_uint64) for n in (
            "query_chunks", "lists", "units_rep", "units_value", "seed_hits", "chained_hits",
            "gapped_extensions", "gapped_cells", "results", "sequential_build_chunks")] + [
        ("gapped_kernel_ms", c_double), ("seed_rep_kernel_ms", c_double),
        ("seed_value_kernel_ms", c_double), ("seed_rep_launches", c_uint64),
        ("seed_value_launches", c_uint64), ("gapped_launches", c_uint64),
        ("schedule_kernel_ms", c_double), ("traceback_kernel_ms", c_double),
        ("traceback_host_jobs", c_uint64), ("gated_jobs", c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class QueryListsC(Structure):
    _fields_ = [("number_lists", c_uint64), ("number_entries", c_uint64), ("hash", c_void_p),
                ("entry_offsets", c_void_p), ("entry_positions", c_void_p),
                ("entry_queries", c_void_p), ("entry_checks", c_void_p), ("seconds", c_double)]


class GappedTrace(Structure):
    _fields_ = [("query_position", c_uint32), ("database_position", c_uint32),
                ("chunk", c_uint32), ("reverse", c_uint32), ("cells", c_uint32)]


SCHEDULE_PLAIN, SCHEDULE_LENGTH_SORTED, SCHEDULE_LEARNED = 0, 1, 2

_SIGNATURES.update({
    "gz_queries_create": (c_int, [c_void_p, c_char_p, c_void_p, c_uint64,
                                  POINTER(QueryOptions), POINTER(c_void_p)]),
    "gz_queries_read_fasta": (c_int, [c_char_p, POINTER(QueryOptions), POINTER(c_void_p)]),
    "gz_queries_destroy": (c_int, [c_void_p]),
    "gz_queries_view": (c_int, [c_void_p, POINTER(QueriesView)]),
    "gz_queries_cutoffs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_int)]),
    "gz_query_lists_build": (c_int, [c_void_p, c_void_p, c_int, POINTER(QueryListsC)]),
    "gz_query_lists_free": (None, [POINTER(QueryListsC)]),
    "gz_database_create": (c_int, [c_void_p, c_char_p, c_void_p, c_uint64,
                                   POINTER(DatabaseOptions), POINTER(c_void_p)]),
    "gz_database_read_fasta": (c_int, [c_char_p, POINTER(DatabaseOptions), POINTER(c_void_p)]),
    "gz_database_destroy": (c_int, [c_void_p]),
    "gz_database_load": (c_int, [c_char_p, POINTER(c_void_p)]),
    "gz_database_save": (c_int, [c_void_p, c_char_p]),
    "gz_database_info": (c_int, [c_void_p, POINTER(DatabaseInfo)]),
    "gz_database_chunk": (c_int, [c_void_p, c_uint32, POINTER(ChunkView)]),
    "gz_aligner_default_parameters": (None, [POINTER(AlignerParameters)]),
    "gz_aligner_create": (c_int, [POINTER(AlignerParameters), POINTER(c_void_p)]),
    "gz_aligner_destroy": (c_int, [c_void_p]),
    "gz_aligner_set_database": (c_int, [c_void_p, c_void_p]),
    "gz_aligner_set_cost_model": (c_int, [c_void_p, c_void_p, c_size_t]),
    "gz_aligner_tuner_state": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "gz_aligner_graph_counts": (c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64)]),
    "gz_aligner_align": (c_int, [c_void_p, c_void_p, POINTER(c_void_p), POINTER(c_size_t),
                                 POINTER(AlignStats)]),
    "gz_aligner_traces": (c_uint64, [c_void_p, c_void_p, c_uint64]),
    "gz_aligner_last_concatenated_queries": (c_int, [c_void_p, POINTER(c_void_p),
                                                     POINTER(c_uint64)]),
    "gz_align_fasta": (c_int, [c_char_p, c_char_p, c_char_p, c_int,
                               POINTER(AlignerParameters), POINTER(DatabaseOptions)]),
    "gz_align_main": (c_int, [c_char_p, c_char_p, c_char_p, c_int, POINTER(AlignerParameters)]),
    "gz_free": (None, [c_void_p]),
    "gz_build_results": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_uint32, c_uint32,
                                 c_void_p, c_uint32, c_int, c_int, c_int, c_void_p, c_void_p]),
    "gz_build_result_host": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, c_int,
                                     c_void_p]),
    "gz_mlp_weight_count": (c_size_t, []),
    "gz_mlp_init_weights": (c_int, [c_uint64, c_void_p]),
    "gz_mlp_create": (c_int, [c_void_p, c_size_t, POINTER(c_void_p)]),
    "gz_mlp_destroy": (c_int, [c_void_p]),
    "gz_mlp_get_weights": (c_int, [c_void_p, c_void_p]),
    "gz_mlp_predict": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, POINTER(c_float)]),
    "gz_mlp_predict_swapped": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "gz_mlp_train_step": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_float,
                                  POINTER(c_float), c_void_p]),
    "gz_mlp_train_epochs": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_size_t,
                                    c_size_t, c_float, POINTER(c_float), POINTER(c_float)]),
    "gz_cost_features": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_int,
                                c_void_p]),
    "gz_cost_features_packed_async": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                              c_int, c_void_p, c_void_p]),
    "gz_seg_mask": (c_int, [c_char_p, c_void_p]),
    "gz_seg_ln_factorial": (c_double, [c_int]),
    "gz_translate": (c_int, [c_char_p, c_void_p, c_void_p]),
    "gz_karlin_block": (None, [c_void_p, c_int, c_int, POINTER(c_float), POINTER(c_float),
                               POINTER(c_float)]),
    "gz_query_karlin": (c_int, [c_void_p, c_uint32, POINTER(c_float), POINTER(c_float),
                                POINTER(c_float)]),
    "gz_search_space": (c_uint64, [c_uint64, c_uint64, c_uint64]),
})


class SeedHits(Structure):
    _fields_ = [("number_hits", c_uint64), ("number_queries", c_uint32),
                ("query_offsets", c_void_p), ("query_positions", c_void_p),
                ("database_positions", c_void_p)]


_SIGNATURES.update({
    "gz_host_seeds_memory_create": (c_int, [c_size_t, POINTER(c_void_p)]),
    "gz_host_seeds_memory_destroy": (c_int, [c_void_p]),
    "gz_host_seeds_memory_size": (c_size_t, [c_void_p]),
    "gz_host_seeds_memory_query_seed_ids": (c_void_p, [c_void_p]),
    "gz_host_seeds_memory_query_seed_starts": (c_void_p, [c_void_p]),
    "gz_host_seeds_memory_query_ids": (c_void_p, [c_void_p]),
    "gz_host_seeds_memory_query_concatenated_positions": (c_void_p, [c_void_p]),
    "gz_host_seeds_memory_database_positions": (c_void_p, [c_void_p]),
    "gz_host_seeds_memory_result_values": (c_void_p, [c_void_p]),
    "gz_stream_controller_create": (c_int, [POINTER(c_void_p)]),
    "gz_stream_controller_destroy": (c_int, [c_void_p]),
    "gz_stream_controller_init": (c_int, [c_void_p, c_size_t]),
    "gz_stream_controller_set_distance_calculator": (c_int, [c_void_p, c_void_p]),
    "gz_stream_controller_set_ungapped_extender": (c_int, [c_void_p, c_void_p]),
    "gz_stream_controller_set_gapped_extender": (c_int, [c_void_p, c_void_p]),
    "gz_stream_controller_run_distance_calculation": (c_int, [c_void_p, c_size_t, c_void_p,
                                                              POINTER(c_int)]),
    "gz_stream_controller_run_ungapped_extension_with_trigger": (
        c_int, [c_void_p, c_size_t, c_size_t, c_void_p, POINTER(c_int)]),
    "gz_stream_controller_run_gapped_extension": (
        c_int, [c_void_p, c_size_t, c_size_t, c_size_t, c_size_t, c_void_p, POINTER(c_int)]),
    "gz_stream_controller_wait_run": (c_int, [c_void_p, c_int]),
    "gz_stream_controller_wait_run_all": (c_int, [c_void_p]),
    "gz_stream_controller_finished_run": (c_int, [c_void_p, c_int]),
    "gz_stream_controller_release_resource": (c_int, [c_void_p, c_int]),
    "gz_stream_controller_available_resources": (c_uint32, [c_void_p]),
    "gz_seed_searcher_create": (c_int, [c_void_p, POINTER(c_void_p)]),
    "gz_seed_searcher_destroy": (c_int, [c_void_p]),
    "gz_seed_searcher_reset": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, POINTER(c_uint32)]),
    "gz_seed_searcher_search_seeds_from_representatives_for_similarity_filtering": (
        c_int, [c_void_p, c_uint32, c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
    "gz_seed_searcher_search_seeds_from_clustering_subsequence": (
        c_int, [c_void_p, c_uint32, c_uint32, POINTER(SeedHits)]),
    "gz_seed_searcher_search_seeds_from_non_clustering_subsequence": (
        c_int, [c_void_p, c_uint32, c_uint32, POINTER(SeedHits)]),
    "gz_seed_searcher_search_seeds": (c_int, [c_void_p, c_uint32, c_uint32, POINTER(SeedHits)]),
})


def _names_array(names):
    arr = (c_char_p * len(names))(*[n.encode() if isinstance(n, str) else n for n in names])
    return arr


def _read_array(addr, dtype, count):
    if count == 0 or not addr:
        return np.zeros(0, dtype)
    n = int(count) * np.dtype(dtype).itemsize
    return np.frombuffer(ctypes.string_at(addr, n), dtype).copy()


class Queries:
    """Host-side query set (Queries): encoding, SEG, translation, Karlin."""

    def __init__(self, names, residues, offsets, translated=False, filter=True, threads=0):
        offs = np.ascontiguousarray(offsets, np.uint64)
        o = QueryOptions(int(translated), int(filter), int(threads))
        h = c_void_p()
        nm = _names_array(names)
        check(lib().gz_queries_create(nm, residues, ptr(offs), len(names), ctypes.byref(o),
                                      ctypes.byref(h)))
        self.handle = h
        self.translated = bool(translated)

    @classmethod
    def from_fasta(cls, path, translated=False, filter=True, threads=0):
        self = cls.__new__(cls)
        o = QueryOptions(int(translated), int(filter), int(threads))
        h = c_void_p()
        check(lib().gz_queries_read_fasta(path.encode(), ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        self.translated = bool(translated)
        return self

    def view(self):
        v = QueriesView()
        check(lib().gz_queries_view(self.handle, ctypes.byref(v)))
        n = v.number_queries
        offsets = _read_array(v.offsets, np.uint64, n + 1)
        out = {
            "number_queries": n,
            "codes": _read_array(v.codes, np.uint8, int(offsets[-1]) if n else 0),
            "offsets": offsets,
            "lambda": _read_array(v.lambda_, np.float32, n),
            "K": _read_array(v.K, np.float32, n),
            "translated": v.translated,
            "dna_lengths": _read_array(v.dna_lengths, np.uint32, n) if v.translated else None,
        }
        names_p = ctypes.cast(v.names, POINTER(c_char_p))
        out["names"] = [names_p[i] for i in range(n)]
        return out

    def __del__(self):
        try:
            if self.handle:
                lib().gz_queries_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Database:
    """In-memory database (Database + SeedSearcherDatabaseParameters index)."""

    def __init__(self, names, residues, offsets, clustering=True, subsequence_length=10,
                 seed_threshold=39, chunk_size=0, threads=0):
        offs = np.ascontiguousarray(offsets, np.uint64)
        o = DatabaseOptions(int(clustering), subsequence_length, seed_threshold, chunk_size,
                            threads)
        h = c_void_p()
        check(lib().gz_database_create(_names_array(names), residues, ptr(offs), len(names),
                                       ctypes.byref(o), ctypes.byref(h)))
        self.handle = h

    @classmethod
    def from_fasta(cls, path, clustering=True, subsequence_length=10, seed_threshold=39,
                   chunk_size=0, threads=0):
        self = cls.__new__(cls)
        o = DatabaseOptions(int(clustering), subsequence_length, seed_threshold, chunk_size,
                            threads)
        h = c_void_p()
        check(lib().gz_database_read_fasta(path.encode(), ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        return self

    @classmethod
    def load(cls, prefix):
        """Database::Load: <prefix>.inf and <prefix>_<i>.{inf,off,nam,seq,scp,sdp}."""
        self = cls.__new__(cls)
        h = c_void_p()
        check(lib().gz_database_load(prefix.encode(), ctypes.byref(h)))
        self.handle = h
        return self

    def save(self, prefix):
        """Database::Save: the same file set."""
        check(lib().gz_database_save(self.handle, prefix.encode()))

    def info(self):
        i = DatabaseInfo()
        check(lib().gz_database_info(self.handle, ctypes.byref(i)))
        return i

    def chunk(self, c):
        v = ChunkView()
        check(lib().gz_database_chunk(self.handle, c, ctypes.byref(v)))
        nh = v.number_hashes
        out = {
            "sequence": _read_array(v.sequence, np.uint8, v.sequence_length),
            "number_sequences": v.number_sequences,
            "offsets": _read_array(v.offsets, np.uint32, v.number_sequences + 1),
            "hashes": _read_array(v.hashes, np.uint32, nh),
            "rep_offsets": _read_array(v.rep_offsets, np.uint32, nh + 1),
            "rep_positions": _read_array(v.rep_positions, np.uint32, v.number_reps),
            "rep_member_offsets": _read_array(v.rep_member_offsets, np.uint32,
                                              v.number_reps + 1),
            "member_positions": _read_array(v.member_positions, np.uint32, v.number_members),
            "value_offsets": _read_array(v.value_offsets, np.uint32, nh + 1),
            "value_positions": _read_array(v.value_positions, np.uint32, v.number_values),
            "clustering": v.clustering,
        }
        names_p = ctypes.cast(v.names, POINTER(c_char_p))
        out["names"] = [names_p[i] for i in range(v.number_sequences)]
        return out

    def chunk_sequence(self, c):
        """The codes of chunk c alone (without copying its index arrays)."""
        v = ChunkView()
        check(lib().gz_database_chunk(self.handle, c, ctypes.byref(v)))
        return _read_array(v.sequence, np.uint8, v.sequence_length)

    def __del__(self):
        try:
            if self.handle:
                lib().gz_database_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def default_parameters(**kw):
    p = AlignerParameters()
    lib().gz_aligner_default_parameters(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Aligner:
    """AlignerGpu on one device: presearch on the GPU, results on the host."""

    def __init__(self, database, **params):
        self.params = default_parameters(**params)
        h = c_void_p()
        check(lib().gz_aligner_create(ctypes.byref(self.params), ctypes.byref(h)))
        self.handle = h
        self.database = database
        check(lib().gz_aligner_set_database(h, database.handle))

    def set_cost_model(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        check(lib().gz_aligner_set_cost_model(self.handle, ptr(flat), flat.size))

    def align(self, queries):
        """Output text and stats; `last_call_seconds` is the wall time of the
        C call alone (the copy into a Python bytes object excluded)."""
        out = c_void_p()
        n = c_size_t()
        st = AlignStats()
        t0 = time.perf_counter()
        check(lib().gz_aligner_align(self.handle, queries.handle, ctypes.byref(out),
                                     ctypes.byref(n), ctypes.byref(st)))
        self.last_call_seconds = time.perf_counter() - t0
        text = ctypes.string_at(out, n.value) if n.value else b""
        lib().gz_free(out)
        return text, st.as_dict()

    def graph_counts(self):
        """(graphs instantiated, launches that updated a graph) of the batched form."""
        a, b = c_uint64(), c_uint64()
        check(lib().gz_aligner_graph_counts(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def tuner_state(self):
        """The batched form's autotuner: chosen configuration per query-length
        bucket (-1: not seen), ms per job and trials per (bucket, config)."""
        chosen = np.zeros(4, np.int32)
        ms = np.zeros(16, np.float64)
        trials = np.zeros(16, np.uint32)
        check(lib().gz_aligner_tuner_state(self.handle, ptr(chosen), ptr(ms), ptr(trials)))
        return chosen, ms.reshape(4, 4), trials.reshape(4, 4)

    def traces(self):
        n = lib().gz_aligner_traces(self.handle, None, 0)
        arr = (GappedTrace * max(int(n), 1))()
        n = lib().gz_aligner_traces(self.handle, arr, n)
        a = np.frombuffer(bytes(arr), np.uint32).reshape(-1, 5)[:n]
        return a.copy()

    def last_concatenated_queries(self):
        p = c_void_p()
        n = c_uint64()
        check(lib().gz_aligner_last_concatenated_queries(self.handle, ctypes.byref(p),
                                                         ctypes.byref(n)))
        return _read_array(p.value, np.uint8, n.value)

    def __del__(self):
        try:
            if self.handle:
                lib().gz_aligner_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Mlp:
    """The learned extension-cost MLP (bf16 MFMA kernels, fp32 master weights)."""

    def __init__(self, flat=None, seed=1):
        if flat is None:
            flat = self.init_weights(seed)
        flat = np.ascontiguousarray(flat, np.float32)
        h = c_void_p()
        check(lib().gz_mlp_create(ptr(flat), flat.size, ctypes.byref(h)))
        self.handle = h

    @staticmethod
    def weight_count():
        return int(lib().gz_mlp_weight_count())

    @staticmethod
    def init_weights(seed):
        w = np.zeros(Mlp.weight_count(), np.float32)
        check(lib().gz_mlp_init_weights(seed, ptr(w)))
        return w

    def weights(self):
        w = np.zeros(self.weight_count(), np.float32)
        check(lib().gz_mlp_get_weights(self.handle, ptr(w)))
        return w

    def predict(self, codes):
        codes = np.ascontiguousarray(codes, np.uint8).reshape(-1, 128)
        out = np.zeros(len(codes), np.float32)
        ms = c_float()
        check(lib().gz_mlp_predict(self.handle, ptr(codes), len(codes), ptr(out),
                                   ctypes.byref(ms)))
        self.last_kernel_ms = ms.value
        return out

    def train_step(self, codes, targets, lr, want_grad=False):
        codes = np.ascontiguousarray(codes, np.uint8).reshape(-1, 128)
        t = np.ascontiguousarray(targets, np.float32)
        loss = c_float()
        g = np.zeros(self.weight_count(), np.float32) if want_grad else None
        check(lib().gz_mlp_train_step(self.handle, ptr(codes), ptr(t), len(codes), lr,
                                      ctypes.byref(loss), ptr(g)))
        return (loss.value, g) if want_grad else loss.value

    def train_epochs(self, codes, targets, order, batch, lr):
        """Every whole batch of `order` (row indices into codes / targets) as
        one training step, the set resident on the device.  Returns (mean
        loss, device milliseconds)."""
        codes = np.ascontiguousarray(codes, np.uint8).reshape(-1, 128)
        t = np.ascontiguousarray(targets, np.float32)
        o = np.ascontiguousarray(order, np.uint32)
        loss, ms = c_float(), c_float()
        check(lib().gz_mlp_train_epochs(self.handle, ptr(codes), ptr(t), len(codes), ptr(o),
                                        o.size, batch, lr, ctypes.byref(loss),
                                        ctypes.byref(ms)))
        return loss.value, ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().gz_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def query_lists(queries, database, on_device):
    """Hash lists of the query set taken as one chunk (device or host build):
    dict of numpy arrays hash, entry_offsets, entry_positions (in the
    shared-delimiter concatenation), entry_queries, entry_checks, plus the
    build time in seconds."""
    ql = QueryListsC()
    check(lib().gz_query_lists_build(queries.handle, database.handle, int(on_device),
                                     ctypes.byref(ql)))
    try:
        nl, ne = int(ql.number_lists), int(ql.number_entries)

        def take(p, n, ty):
            return np.frombuffer(ctypes.string_at(p, n * np.dtype(ty).itemsize), ty).copy() \
                if n else np.zeros(0, ty)
        out = {"hash": take(ql.hash, nl, np.uint32),
               "entry_offsets": take(ql.entry_offsets, nl + 1, np.uint32),
               "entry_positions": take(ql.entry_positions, ne, np.uint32),
               "entry_queries": take(ql.entry_queries, ne, np.uint32),
               "entry_checks": take(ql.entry_checks, ne, np.uint8),
               "seconds": float(ql.seconds)}
    finally:
        lib().gz_query_lists_free(ctypes.byref(ql))
    return out


def cost_features(queries, database, qpos, dpos, reverse):
    q = np.ascontiguousarray(queries, np.uint8)
    d = np.ascontiguousarray(database, np.uint8)
    qp = np.ascontiguousarray(qpos, np.uint32)
    dp = np.ascontiguousarray(dpos, np.uint32)
    out = np.zeros((len(qp), 128), np.uint8)
    lib().gz_cost_features(ptr(q), ptr(d), ptr(qp), ptr(dp), len(qp), int(reverse), ptr(out))
    return out
