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
        _ref = L
    return _ref


# ---- thin wrappers ------------------------------------------------------------
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
