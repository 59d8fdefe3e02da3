"""The C ABI: every function declared in include/ghostz_hip.h is exported by
libghostz_amd.so, and the ctypes structures mirror the C layouts.  CPU only
(no compute calls)."""
import ctypes
import os
import re

import gz_native as gz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ghostz_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(gz_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_t")))


def test_every_declared_function_is_exported():
    lib = ctypes.CDLL(gz.LIB_PATH)
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    names = set(declared_functions())
    unbound = sorted(names - set(gz._SIGNATURES))
    # every entry point is reachable from Python with a declared signature
    assert not unbound, unbound


def test_struct_sizes_match_c():
    # sizes fixed by the C declarations (x86-64, natural alignment)
    assert ctypes.sizeof(gz.QueryOptions) == 12
    assert ctypes.sizeof(gz.QueriesView) == 64
    assert ctypes.sizeof(gz.DatabaseOptions) == 32
    assert ctypes.sizeof(gz.DatabaseInfo) == 168
    assert ctypes.sizeof(gz.ChunkView) == 128
    assert ctypes.sizeof(gz.AlignerParameters) == 64
    assert ctypes.sizeof(gz.GappedTrace) == 20
    assert ctypes.sizeof(gz.AlignStats) == 232


def test_errors_are_reported_not_fatal():
    lib = gz.lib()
    assert lib.gz_aligner_align(None, None, None, None, None) != 0
    assert b"null" in lib.gz_last_error()
    assert lib.gz_mlp_create(None, 0, None) != 0
