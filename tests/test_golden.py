"""The oracle and the product's host code against the committed golden
fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
reference's own leaf code).  These run without the reference tree."""
import ctypes
import os

import numpy as np

import gz_native as gz
import oracle_py

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DELIM = 25


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_gapped_extender_fixture_oracle():
    g = load("gapped_extender.npz")
    m, nl = gz.blosum62()
    q, d = g["query"], g["database"]
    off = g["edit_offsets"]
    for k in range(g["qstart"].size):
        s, p0, p1, ed, _ = oracle_py.gapped_one_side(
            "oracle", q, int(g["qstart"][k]), int(g["seq0_len"][k]), d, int(g["dstart"][k]),
            DELIM, int(g["reverse"][k]), m, nl, int(g["gap_open"][k]), int(g["gap_ext"][k]),
            int(g["cutoff"][k]), traceback=True)
        assert (s, p0, p1) == (g["score"][k], g["p0"][k], g["p1"][k]), k
        assert ed == g["edits"][off[k]:off[k + 1]].tolist(), k


def test_ungapped_distance_hash_fixture_oracle():
    g = load("ungapped_distance_hash.npz")
    O = oracle_py.oracle()
    m, nl = gz.blosum62()
    q, d = g["query"], g["database"]
    for a, b, rev, cut, s, b0, b1 in g["ungapped"]:
        x0, x1 = ctypes.c_int(), ctypes.c_int()
        got = O.gzo_ungapped_one_side_positions(
            ctypes.c_void_p(q.ctypes.data + int(a)), ctypes.c_void_p(d.ctypes.data + int(b)),
            DELIM, int(rev), oracle_py.ptr(m), nl, int(cut), ctypes.byref(x0), ctypes.byref(x1))
        assert (got, x0.value, x1.value) == (s, b0, b1)
    rmap = g["reduced_map"]
    for a, b, dist in g["distance"]:
        assert O.gzo_distance(oracle_py.ptr(q), int(a), oracle_py.ptr(d), int(b), 10,
                              oracle_py.ptr(rmap)) == dist
    # the oracle's hash takes the +1 shifted map and scores (code 0 = delimiter)
    mx = int(g["max_regular"])
    shifted_map = (rmap.astype(np.int64) + 1).astype(np.uint8)
    shifted_scores = np.concatenate([[0], g["hash_scores"]]).astype(np.int32)
    seq = g["hash_sequence"]
    for i, ret, h, hl in g["hash"]:
        h1, l1 = ctypes.c_uint32(), ctypes.c_uint32()
        r = O.gzo_hash(ctypes.c_void_p(seq.ctypes.data + int(i)), oracle_py.ptr(shifted_map), 25,
                       oracle_py.ptr(shifted_scores), mx + 1, 39, 8, 4, ctypes.byref(h1),
                       ctypes.byref(l1))
        assert r == ret
        if r == 0:
            assert (h1.value, l1.value) == (h, hl)


def test_statistics_fixture_product():
    g = load("statistics.npz")
    seqs, offs = g["sequences"], g["offsets"]
    for k in range(len(offs) - 1):
        s = np.ascontiguousarray(seqs[offs[k]:offs[k + 1]], np.uint8)
        lam, K, H = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ok = gz.lib().gz_query_karlin(gz.ptr(s), s.size, ctypes.byref(lam), ctypes.byref(K),
                                      ctypes.byref(H))
        want = g["karlin"][k]
        assert ok == int(want[0])
        if ok:
            assert (np.float32(lam.value), np.float32(K.value), np.float32(H.value)) == tuple(
                np.float32(want[1:]))
    for ql, dl, dn, sp in g["search_space"]:
        assert gz.lib().gz_search_space(int(ql), int(dl), int(dn)) == sp


def test_seg_translate_fixture_product():
    g = load("seg_translate.npz")
    po = g["protein_offsets"]
    for k in range(len(po) - 1):
        s = g["protein"][po[k]:po[k + 1]].tobytes()
        out = ctypes.create_string_buffer(len(s) + 1)
        st = gz.lib().gz_seg_mask(s, out)
        assert st == g["seg_status"][k]
        if st == 0:
            assert out.raw[:len(s)] == g["masked"][po[k]:po[k + 1]].tobytes()
    do, fo = g["dna_offsets"], g["frame_offsets"]
    for k in range(len(do) - 1):
        s = g["dna"][do[k]:do[k + 1]].tobytes()
        out = ctypes.create_string_buffer(2 * len(s) + 16)
        lens = np.zeros(6, np.uint32)
        gz.lib().gz_translate(s, out, gz.ptr(lens))
        assert np.array_equal(lens, g["frame_lengths"][k])
        t = int(lens.sum())
        assert out.raw[:t] == g["frames"][fo[k]:fo[k + 1]].tobytes()


def subjects_of_chunk(seq):
    """Names, residues and offsets of the subjects in a chunk's concatenated
    codes (leading delimiter, subject + delimiter ...), decoded to letters."""
    t = gz.encode_table()
    letters = np.zeros(32, np.uint8)
    for c in range(ord("A"), ord("Z") + 1):
        if t[c] < 25 and letters[t[c]] == 0:
            letters[t[c]] = c
    letters[t[ord("*")]] = ord("*")
    cuts = np.flatnonzero(seq == DELIM)
    parts = [seq[a + 1:b] for a, b in zip(cuts[:-1], cuts[1:])]
    residues = letters[np.concatenate(parts)].tobytes()
    offsets = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint64)
    return [f"s{k}" for k in range(len(parts))], residues, offsets


def test_database_index_fixture_product():
    g = load("db_index.npz")
    dn, dr, do = subjects_of_chunk(g["sequence"])
    for clustering in (1, 0):
        db = gz.Database(dn, dr, do, clustering=bool(clustering))
        ch = db.chunk(0)
        assert np.array_equal(ch["sequence"], g["sequence"])
        for k in ("hashes", "value_offsets", "value_positions"):
            assert np.array_equal(ch[k], g[f"c{clustering}_{k}"]), (clustering, k)
        if clustering:
            for k in ("rep_offsets", "rep_positions", "rep_member_offsets", "member_positions"):
                assert np.array_equal(ch[k], g[f"c1_{k}"]), k
