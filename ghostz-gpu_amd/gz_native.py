"""ctypes binding of libghostz_amd.so (the C ABI declared in include/ghostz_hip.h).

The library is the product: HIP kernels for gfx950 plus the C++ host side.
Loading fails loudly when it has not been built -- there is no fallback.
"""
import ctypes
import os
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
