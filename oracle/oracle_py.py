"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
driver (oracle/_ref/libghostz_ref.so, present only where the reference tree was
available at build time).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never imports this module.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint8, c_uint32, \
    c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libghostz_ref.so")

_oracle = None
_ref = None


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: build with `make -C oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        L.gzo_ungapped_one_side.restype = c_int
        L.gzo_ungapped_one_side.argtypes = [c_void_p, c_void_p, c_uint8, c_int, c_void_p,
                                            c_uint32, c_int]
        L.gzo_ungapped_one_side_positions.restype = c_int
        L.gzo_ungapped_one_side_positions.argtypes = [c_void_p, c_void_p, c_uint8, c_int,
                                                      c_void_p, c_uint32, c_int,
                                                      POINTER(c_int), POINTER(c_int)]
        L.gzo_ungapped_with_trigger.restype = c_int
        L.gzo_ungapped_with_trigger.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_uint8,
                                                c_void_p, c_uint32, c_int, c_int]
        L.gzo_distance.restype = c_int
        L.gzo_distance.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, c_void_p]
        L.gzo_build_chunk.restype = c_int
        L.gzo_build_chunk.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32,
                                      c_uint8, c_void_p, c_uint32, c_int, c_int, c_int,
                                      c_void_p]
        L.gzo_gapped_one_side.restype = c_int
        L.gzo_gapped_one_side.argtypes = [c_void_p, c_uint32, c_void_p, c_uint8, c_int,
                                          c_void_p, c_uint32, c_int, c_int, c_int,
                                          POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                          c_void_p, POINTER(c_uint32), POINTER(c_uint64)]
        L.gzo_chain_filter.restype = c_uint32
        L.gzo_chain_filter.argtypes = [c_void_p, c_uint32, c_int, c_void_p, c_uint8, c_void_p,
                                       c_uint32, c_void_p, c_uint32]
        L.gzo_hash.restype = c_int
        L.gzo_hash.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_int,
                               c_uint32, c_uint32, POINTER(c_uint32), POINTER(c_uint32)]
        L.gzo_align.restype = c_int
        L.gzo_align.argtypes = [c_void_p, c_void_p, c_void_p, POINTER(c_void_p),
                                POINTER(c_size_t)]
        L.gzo_free.restype = None
        L.gzo_free.argtypes = [c_void_p]
        L.gzo_query_lists.restype = c_int
        L.gzo_query_lists.argtypes = [c_void_p, c_void_p, POINTER(c_uint64), POINTER(c_void_p),
                                      POINTER(c_void_p), POINTER(c_uint64), POINTER(c_void_p),
                                      POINTER(c_void_p), POINTER(c_void_p)]
        L.gzo_last_stats.restype = None
        L.gzo_last_stats.argtypes = [c_void_p]
        _oracle = L
    return _oracle


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own leaf code (only where oracle/_ref was built)."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_PATH)
        L.gzr_read_matrix.argtypes = [c_char_p, c_void_p, POINTER(c_uint32)]
        L.gzr_protein_encode_table.argtypes = [c_void_p]
        L.gzr_reduced_map.argtypes = [c_char_p, c_void_p, POINTER(c_uint32)]
        L.gzr_ungapped_one_side.restype = c_int
        L.gzr_ungapped_one_side.argtypes = [c_void_p, c_void_p, c_uint8, c_int, c_void_p,
                                            c_uint32, c_int]
        L.gzr_ungapped_one_side_positions.restype = c_int
        L.gzr_ungapped_one_side_positions.argtypes = [c_void_p, c_void_p, c_uint8, c_int,
                                                      c_void_p, c_uint32, c_int,
                                                      POINTER(c_int), POINTER(c_int)]
        L.gzr_gapped_one_side.argtypes = [c_void_p, c_uint32, c_void_p, c_uint8, c_int,
                                          c_void_p, c_uint32, c_int, c_int, c_int,
                                          POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                          c_void_p, POINTER(c_uint32)]
        L.gzr_distance.restype = c_int
        L.gzr_distance.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32, c_uint32, c_void_p]
        L.gzr_hash.restype = c_int
        L.gzr_hash.argtypes = [c_void_p, c_void_p, c_uint32, c_int, c_uint32, c_void_p,
                               POINTER(c_uint32), POINTER(c_uint32)]
        L.gzr_karlin_blk.argtypes = [c_void_p, c_int, c_int, POINTER(c_float),
                                     POINTER(c_float), POINTER(c_float)]
        L.gzr_query_karlin.restype = c_int
        L.gzr_query_karlin.argtypes = [c_void_p, c_uint32, c_void_p, c_uint32,
                                       POINTER(c_float), POINTER(c_float), POINTER(c_float)]
        L.gzr_ideal_karlin.argtypes = [c_void_p, c_uint32, POINTER(c_float), POINTER(c_float),
                                       POINTER(c_float)]
        L.gzr_normalized2nominal.restype = c_int
        L.gzr_normalized2nominal.argtypes = [c_float, c_float, c_float]
        L.gzr_normalized_cutoff2nominal.restype = c_int
        L.gzr_normalized_cutoff2nominal.argtypes = [c_float, c_float, c_float]
        L.gzr_nominal2normalized.restype = c_float
        L.gzr_nominal2normalized.argtypes = [c_int, c_float, c_float]
        L.gzr_nominal2evalue.restype = c_double
        L.gzr_nominal2evalue.argtypes = [c_int, c_uint64, c_float, c_float]
        L.gzr_search_space.restype = c_uint64
        L.gzr_search_space.argtypes = [c_uint64, c_uint64, c_uint64]
        L.gzr_translate.restype = c_int
        L.gzr_translate.argtypes = [c_char_p, c_void_p, c_void_p]
        L.gzr_seg.restype = c_int
        L.gzr_seg.argtypes = [c_char_p, c_void_p]
        L.gzr_build_index.restype = c_void_p
        L.gzr_build_index.argtypes = [c_void_p, c_uint32, c_int, c_uint32, c_int, c_void_p,
                                      c_uint32, c_char_p]
        L.gzr_index_sizes.argtypes = [c_void_p, c_void_p]
        L.gzr_index_copy.argtypes = [c_void_p, c_void_p]
        L.gzr_index_free.argtypes = [c_void_p]
        L.gzr_index_files.restype = None
        L.gzr_index_files.argtypes = [c_void_p, c_uint32, c_int, c_uint32, c_int, c_void_p,
                                      c_uint32, c_char_p, POINTER(c_void_p), POINTER(c_uint64),
                                      POINTER(c_void_p), POINTER(c_uint64)]
        L.gzr_free_bytes.argtypes = [c_void_p]
        _ref = L
    return _ref


# ---- thin wrappers ------------------------------------------------------------
def ref_index_files(seq, clustering, L, seed_threshold, matrix, nl,
                    sets=b"A KR EDNQ C G H ILVM FYW P ST"):
    """The reference's own .sdp and .scp bytes for one concatenated chunk
    (SeedSearcherDatabaseParameters::Save / SaveCommonParameters)."""
    R = ref()
    seq = np.ascontiguousarray(seq, np.uint8)
    m = np.ascontiguousarray(matrix, np.int32)
    a, b = c_void_p(), c_void_p()
    na, nb = c_uint64(), c_uint64()
    R.gzr_index_files(ptr(seq), seq.size, int(clustering), L, seed_threshold, ptr(m), nl, sets,
                      ctypes.byref(a), ctypes.byref(na), ctypes.byref(b), ctypes.byref(nb))
    sdp = ctypes.string_at(a, na.value)
    scp = ctypes.string_at(b, nb.value)
    R.gzr_free_bytes(a)
    R.gzr_free_bytes(b)
    return sdp, scp


def write_reference_database(prefix, chunks, max_length, total_length, total_sequences,
                             clustering, L, seed_threshold, matrix, nl):
    """Writes a database in the reference's file set: the chunk files that
    DatabaseChunk::Save{Infomation,Offsets,Names,ConcatenatedSequence} write
    (raw fields, src/database_chunk.h:403-447; one name per line), the
    reference's own .sdp / .scp streams, and Database::SaveInformation's
    <prefix>.inf (src/database.h:540-554).  chunks: (sequence codes, offsets,
    names) per chunk."""
    import struct
    for i, (seq, offsets, names) in enumerate(chunks):
        p = f"{prefix}_{i}"
        seq = np.ascontiguousarray(seq, np.uint8)
        with open(p + ".inf", "wb") as f:
            f.write(struct.pack("<II", len(names), seq.size))
        with open(p + ".off", "wb") as f:
            f.write(np.ascontiguousarray(offsets, np.uint32).tobytes())
        with open(p + ".nam", "w") as f:
            for n in names:
                f.write((n.decode() if isinstance(n, bytes) else n) + "\n")
        with open(p + ".seq", "wb") as f:
            f.write(seq.tobytes())
        sdp, scp = ref_index_files(seq, clustering, L, seed_threshold, matrix, nl)
        with open(p + ".sdp", "wb") as f:
            f.write(sdp)
        with open(p + ".scp", "wb") as f:
            f.write(scp)
    with open(prefix + ".inf", "wb") as f:
        f.write(struct.pack("<iIQQB", len(chunks), max_length, total_length, total_sequences, 25))
def gapped_one_side(lib_kind, seq, start, seq0_len, db, dstart, delim, reversed_, matrix, nl,
                    gap_open, gap_ext, cutoff, traceback=False):
    """Runs the oracle ('oracle') or the reference ('ref') gapped extender.
    seq/db are uint8 arrays; start/dstart index the first compared code."""
    s = c_int()
    p0 = c_int()
    p1 = c_int()
    n = c_uint32()
    edits = np.zeros(seq.size + db.size + 8, np.uint8) if traceback else None
    a = ctypes.c_void_p(seq.ctypes.data + int(start))
    b = ctypes.c_void_p(db.ctypes.data + int(dstart))
    seq0_len, delim = int(seq0_len), int(delim)
    if lib_kind == "oracle":
        cells = c_uint64()
        oracle().gzo_gapped_one_side(a, seq0_len, b, delim, int(reversed_), ptr(matrix), nl,
                                     gap_open, gap_ext, cutoff, ctypes.byref(s), ctypes.byref(p0),
                                     ctypes.byref(p1), ptr(edits), ctypes.byref(n),
                                     ctypes.byref(cells))
        extra = cells.value
    else:
        ref().gzr_gapped_one_side(a, seq0_len, b, delim, int(reversed_), ptr(matrix), nl,
                                  gap_open, gap_ext, cutoff, ctypes.byref(s), ctypes.byref(p0),
                                  ctypes.byref(p1), ptr(edits), ctypes.byref(n))
        extra = None
    ed = edits[:n.value].tolist() if traceback else None
    return s.value, p0.value, p1.value, ed, extra


# ---- whole-search oracle (gzo_align) ------------------------------------------
from ctypes import Structure, POINTER as _P  # noqa: E402


class _Chunk(Structure):
    _fields_ = [("sequence", c_void_p), ("sequence_length", c_uint32),
                ("number_sequences", c_uint32), ("offsets", c_void_p), ("names", c_void_p),
                ("number_hashes", c_uint32), ("hashes", c_void_p), ("rep_offsets", c_void_p),
                ("rep_positions", c_void_p), ("rep_member_offsets", c_void_p),
                ("member_positions", c_void_p), ("value_offsets", c_void_p),
                ("value_positions", c_void_p), ("clustering", c_int)]


class _Database(Structure):
    _fields_ = [("number_chunks", c_uint32), ("chunks", c_void_p), ("total_length", c_uint64),
                ("total_sequences", c_uint64), ("subsequence_length", c_uint32),
                ("reduced_code_map", c_void_p), ("hash_code_map", c_void_p),
                ("hash_code_scores", c_void_p), ("hash_max_code", c_uint32),
                ("hash_score_threshold", c_int), ("hash_max_length", c_uint32),
                ("hash_shift_size", c_uint32), ("max_mismatches", c_int)]


class _Queries(Structure):
    _fields_ = [("number_queries", c_uint32), ("sequences", c_void_p), ("offsets", c_void_p),
                ("names", c_void_p), ("lambda_", c_void_p), ("K", c_void_p),
                ("translated", c_int), ("dna_lengths", c_void_p)]


class _Parameters(Structure):
    _fields_ = [("matrix", c_void_p), ("number_letters", c_uint32), ("delimiter", c_uint8),
                ("gap_open", c_int), ("gap_extension", c_int),
                ("normalized_ungapped_cutoff", c_float), ("normalized_gapped_trigger", c_float),
                ("normalized_presearch_cutoff", c_float), ("normalized_result_cutoff", c_float),
                ("max_results", c_uint32), ("max_results_per_subject", c_uint32)]


class _Stats(Structure):
    _fields_ = [(n, c_uint64) for n in ("hash_lists", "distance_calculations",
                                         "ungapped_extensions", "hits",
                                         "hits_after_chain_filter", "gapped_extensions",
                                         "stale_traceback_reads")]


def _structs(queries_view, chunks, info, keep):
    """gzo_queries / gzo_database for prepared inputs (arrays kept alive in
    `keep`)."""
    def arr(a, dtype):
        a = np.ascontiguousarray(a, dtype)
        keep.append(a)
        return a.ctypes.data if a.size else None

    def names(lst):
        a = (c_char_p * max(len(lst), 1))(*lst)
        keep.append(a)
        return ctypes.cast(a, c_void_p)

    cs = (_Chunk * max(len(chunks), 1))()
    keep.append(cs)
    for i, ch in enumerate(chunks):
        cs[i] = _Chunk(arr(ch["sequence"], np.uint8), len(ch["sequence"]),
                       ch["number_sequences"], arr(ch["offsets"], np.uint32), names(ch["names"]),
                       len(ch["hashes"]), arr(ch["hashes"], np.uint32),
                       arr(ch["rep_offsets"], np.uint32), arr(ch["rep_positions"], np.uint32),
                       arr(ch["rep_member_offsets"], np.uint32),
                       arr(ch["member_positions"], np.uint32),
                       arr(ch["value_offsets"], np.uint32),
                       arr(ch["value_positions"], np.uint32), int(ch["clustering"]))
    rmap = np.array(list(info.reduced_code_map), np.uint8)
    hmap = np.array(list(info.hash_code_map), np.uint8)
    hsc = np.array(list(info.hash_code_scores), np.int32)
    db = _Database(len(chunks), ctypes.cast(cs, c_void_p), info.total_length,
                   info.total_sequences, info.subsequence_length, arr(rmap, np.uint8),
                   arr(hmap, np.uint8), arr(hsc, np.int32), info.hash_max_code,
                   info.hash_score_threshold, info.hash_max_length, info.hash_shift,
                   info.max_mismatches)
    qv = queries_view
    q = _Queries(int(qv["number_queries"]), arr(qv["codes"], np.uint8),
                 arr(qv["offsets"], np.uint64), names(qv["names"]),
                 arr(qv["lambda"], np.float32), arr(qv["K"], np.float32),
                 int(qv["translated"]),
                 arr(qv["dna_lengths"], np.uint32) if qv["translated"] else None)
    return q, db, arr


def align(queries_view, chunks, info, matrix, number_letters, gap_open=11, gap_extension=1,
          cutoffs=(7.0, 22.0, 15.0, 25.0), max_results=10, max_results_per_subject=1):
    """Runs the oracle search on already-prepared inputs: queries_view as
    returned by gz_native.Queries.view(), chunks as gz_native.Database.chunk()
    dicts and info a gz_native.DatabaseInfo.  Returns (text bytes, stats dict)."""
    keep = []
    q, db, arr = _structs(queries_view, chunks, info, keep)
    p = _Parameters(arr(matrix, np.int32), number_letters, 25, gap_open, gap_extension,
                    *cutoffs, max_results, max_results_per_subject)
    out = c_void_p()
    n = c_size_t()
    rc = oracle().gzo_align(ctypes.byref(q), ctypes.byref(db), ctypes.byref(p),
                            ctypes.byref(out), ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"gzo_align failed ({rc})")
    text = ctypes.string_at(out, n.value) if n.value else b""
    oracle().gzo_free(out)
    st = _Stats()
    oracle().gzo_last_stats(ctypes.byref(st))
    return text, {k: getattr(st, k) for k, _ in _Stats._fields_}


def query_lists(queries_view, chunks, info):
    """The oracle's hash lists of the queries taken as one chunk, in visiting
    order: dict of hash, entry_offsets, entry_queries, entry_positions
    (query-local) and entry_checks arrays."""
    keep = []
    q, db, _ = _structs(queries_view, chunks, info, keep)
    nl, ne = c_uint64(), c_uint64()
    ptrs = [c_void_p() for _ in range(5)]
    oracle().gzo_query_lists(ctypes.byref(q), ctypes.byref(db), ctypes.byref(nl),
                             ctypes.byref(ptrs[0]), ctypes.byref(ptrs[1]), ctypes.byref(ne),
                             ctypes.byref(ptrs[2]), ctypes.byref(ptrs[3]), ctypes.byref(ptrs[4]))
    sizes = [nl.value, nl.value + 1, ne.value, ne.value, ne.value]
    types = [np.uint32, np.uint32, np.uint32, np.uint32, np.uint8]
    keys = ["hash", "entry_offsets", "entry_queries", "entry_positions", "entry_checks"]
    out = {}
    for k, pp, sz, ty in zip(keys, ptrs, sizes, types):
        buf = ctypes.string_at(pp, sz * np.dtype(ty).itemsize) if sz else b""
        out[k] = np.frombuffer(buf, ty).copy()
        oracle().gzo_free(pp)
    return out


# ---- the product's host flow with emulated device stages (libhostpipe.so) ----
HOSTPIPE_PATH = os.path.join(HERE, "libhostpipe.so")
_hostpipe = None


def hostpipe(path=None):
    global _hostpipe
    if _hostpipe is None or path:
        L = ctypes.CDLL(path or HOSTPIPE_PATH)
        L.gzh_align.restype = c_int
        L.gzh_align.argtypes = [c_void_p, c_char_p, c_void_p, c_uint64, c_int, c_int,
                                c_void_p, c_char_p, c_void_p, c_uint64, c_int, c_uint64,
                                c_void_p, _P(c_void_p), _P(c_size_t), c_void_p]
        L.gzh_free.argtypes = [c_void_p]
        L.gzh_force_stale.restype = None
        L.gzh_force_stale.argtypes = [c_int]
        L.gzh_ungapped_words.restype = None
        L.gzh_ungapped_words.argtypes = [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p,
                                         c_void_p, c_uint32, c_void_p, c_void_p, c_int,
                                         c_void_p]
        L.gzh_distance_words.restype = None
        L.gzh_distance_words.argtypes = [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p,
                                         c_uint32, c_void_p, c_void_p]
        L.gzh_format_g6_check.restype = c_uint64
        L.gzh_format_g6_check.argtypes = [c_void_p, c_uint64, c_void_p, _P(c_uint64)]
        L.gzh_stale_resolution.restype = c_int
        L.gzh_stale_resolution.argtypes = [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32,
                                           c_void_p, c_uint32, c_void_p, c_uint32, c_int, c_int,
                                           c_int, c_void_p, c_void_p]
        if path:
            return L
        _hostpipe = L
    return _hostpipe


def stale_resolution(qcodes, qoff, dcodes, hits, matrix, nl, gap_open, gap_ext, cutoff,
                     lib=None):
    """gzh_stale_resolution: (results read history bytes, resolved, differ from
    the replay, history-dependent) for one chunk; qoff holds each query's
    leading delimiter and, last, the closing delimiter."""
    L = lib or hostpipe()
    qc = np.ascontiguousarray(qcodes, np.uint8)
    qo = np.ascontiguousarray(qoff, np.uint64)
    dc = np.ascontiguousarray(dcodes, np.uint8)
    h = np.ascontiguousarray(hits, np.uint32).ravel()
    m = np.ascontiguousarray(matrix, np.int32)
    out = np.zeros(4, np.uint64)
    res = np.zeros((h.size // 3, 9), np.int32)
    rc = L.gzh_stale_resolution(ptr(qc), ptr(qo), len(qo) - 1, ptr(dc), dc.size, ptr(h),
                                h.size // 3, ptr(m), nl, gap_open, gap_ext, cutoff, ptr(out),
                                ptr(res))
    if rc != 0:
        raise RuntimeError(f"gzh_stale_resolution failed ({rc})")
    return [int(x) for x in out], res


def build_chunk(qcodes, qoff, dcodes, hits, matrix, nl, gap_open, gap_ext, cutoff, delim):
    """The oracle's BuildResults of one chunk (gzo_build_chunk): 9 ints per hit."""
    qc = np.ascontiguousarray(qcodes, np.uint8)
    qo = np.ascontiguousarray(qoff, np.uint64)
    dc = np.ascontiguousarray(dcodes, np.uint8)
    h = np.ascontiguousarray(hits, np.uint32).ravel()
    m = np.ascontiguousarray(matrix, np.int32)
    res = np.zeros((h.size // 3, 9), np.int32)
    oracle().gzo_build_chunk(ptr(qc), ptr(qo), len(qo) - 1, ptr(dc), ptr(h), h.size // 3, delim,
                             ptr(m), nl, gap_open, gap_ext, cutoff, ptr(res))
    return res


def host_align(queries, database, params, translated=False, filter=True, clustering=True,
               db_chunk_size=0, lib=None, force_stale=0):
    """queries/database = (names, residues bytes, offsets); params a
    gz_native.AlignerParameters.  Returns (text, stats dict).  force_stale=n
    makes every n-th emulated device traceback report a stale read."""
    L = lib or hostpipe()
    L.gzh_force_stale(int(force_stale))
    (qn, qr, qo), (dn, dr, do) = queries, database
    qna = (c_char_p * len(qn))(*[n.encode() for n in qn])
    dna = (c_char_p * len(dn))(*[n.encode() for n in dn])
    qo = np.ascontiguousarray(qo, np.uint64)
    do = np.ascontiguousarray(do, np.uint64)
    out = c_void_p()
    n = c_size_t()
    st = np.zeros(6, np.uint64)
    rc = L.gzh_align(ctypes.cast(qna, c_void_p), qr, ptr(qo), len(qn), int(translated),
                     int(filter), ctypes.cast(dna, c_void_p), dr, ptr(do), len(dn),
                     int(clustering), db_chunk_size, ctypes.byref(params), ctypes.byref(out),
                     ctypes.byref(n), ptr(st))
    if rc != 0:
        raise RuntimeError(f"gzh_align failed ({rc})")
    text = ctypes.string_at(out, n.value) if n.value else b""
    L.gzh_free(out)
    keys = ["seed_hits", "chained_hits", "gapped_extensions", "gapped_cells", "results",
            "sequential_build_chunks"]
    return text, dict(zip(keys, (int(x) for x in st)))
