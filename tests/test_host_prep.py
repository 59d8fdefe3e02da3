"""Host preparation of the product against the reference's own leaf code.

oracle/_ref compiles the reference's SequenceSegFilter, Translator,
Statistics, karlin.c and SeedSearcherDatabaseParameters unmodified; every
comparison here is exact (bytes, integers, and float32 values bit for bit).
"""
import ctypes
import math

import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic
from helpers import BACKGROUND

REF = pytest.mark.ref
AMINO = "ARNDCQEGHILKMFPSTWYV"


def _rand_protein_text(rng, n, low_complexity=False):
    if low_complexity and n > 20:
        motif = "".join(rng.choice(list(AMINO), int(rng.integers(1, 4))))
        s = list("".join(rng.choice(list(AMINO), n, p=BACKGROUND)))
        a = int(rng.integers(0, n - 10))
        b = min(n, a + int(rng.integers(10, 60)))
        for i in range(a, b):
            s[i] = motif[(i - a) % len(motif)]
        return "".join(s)
    return "".join(rng.choice(list(AMINO) + ["X", "B", "Z"], n,
                              p=np.concatenate([BACKGROUND * 0.98, [0.01, 0.005, 0.005]])))


@REF
def test_seg_matches_reference():
    rng = np.random.default_rng(11)
    r = oracle_py.ref()
    cases = ["", "A", "ACDEFGHIKL", "ACDEFGHIKLM", "ACDEFGHIKLMN", "AAAAAAAAAAAAAAAAAAAA",
             "QQQQQQQQQQQQQQQQQQQQQQQQQQQQQQQQPPPPPPPPPPPPPPPPPPPPP"]
    for n in list(range(0, 40)) + [50, 100, 300, 1000, 3000]:
        cases.append(_rand_protein_text(rng, n, low_complexity=bool(rng.integers(2))))
    masked_any = 0
    for s in cases:
        a = ctypes.create_string_buffer(len(s) + 1)
        b = ctypes.create_string_buffer(len(s) + 1)
        ra = gz.lib().gz_seg_mask(s.encode(), a)
        rb = r.gzr_seg(s.encode(), b)
        assert ra == rb, s
        if ra == 0:
            assert a.raw[:len(s)] == b.raw[:len(s)], s
            masked_any += a.raw[:len(s)] != s.encode()
    assert masked_any > 3  # the low-complexity cases were masked


def test_seg_ln_factorial_is_rounded_lgamma():
    # the SEG table holds ln(n!) rounded to 6 decimals
    for n in [0, 1, 2, 3, 5, 10, 12, 20, 50, 99, 100]:
        v = gz.lib().gz_seg_ln_factorial(n)
        assert abs(v - round(math.lgamma(n + 1), 6)) < 1e-9, n


@REF
def test_translate_matches_reference():
    rng = np.random.default_rng(12)
    r = oracle_py.ref()
    alphabet = list("ACGT") * 6 + list("acgtNn")
    for n in list(range(0, 30)) + [149, 150, 151, 300, 1001]:
        dna = "".join(rng.choice(alphabet, n)).encode()
        a = ctypes.create_string_buffer(2 * n + 16)
        b = ctypes.create_string_buffer(2 * n + 16)
        la = np.zeros(6, np.uint32)
        lb = np.zeros(6, np.uint32)
        na = gz.lib().gz_translate(dna, a, gz.ptr(la))
        nb = r.gzr_translate(dna, b, oracle_py.ptr(lb))
        assert na == nb and np.array_equal(la, lb), dna
        t = int(la.sum())
        assert a.raw[:t] == b.raw[:t], dna


@REF
def test_karlin_matches_reference():
    rng = np.random.default_rng(13)
    r = oracle_py.ref()
    m, nl = gz.blosum62()
    for trial in range(60):
        n = int(rng.integers(1, 400)) if trial % 5 else int(rng.integers(1, 6))
        body = rng.choice(20, n, p=BACKGROUND).astype(np.uint8)
        if trial % 7 == 0:
            body[:] = body[0]  # degenerate composition
        if trial % 11 == 0:
            body[rng.random(n) < 0.2] = 24
        seq = np.concatenate([[25], body, [25]]).astype(np.uint8)
        la, ka, ha = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        lb, kb, hb = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        oka = gz.lib().gz_query_karlin(gz.ptr(seq), len(seq), ctypes.byref(la),
                                       ctypes.byref(ka), ctypes.byref(ha))
        okb = r.gzr_query_karlin(oracle_py.ptr(seq), len(seq), oracle_py.ptr(m), nl,
                                 ctypes.byref(lb), ctypes.byref(kb), ctypes.byref(hb))
        assert oka == okb
        if oka:
            assert (la.value, ka.value, ha.value) == (lb.value, kb.value, hb.value)
    # karlin.c directly on a few score distributions
    for lo, hi in [(-4, 11), (-1, 1), (-3, 5)]:
        p = rng.random(hi - lo + 1)
        p[-1] = max(p[-1], 0.01)
        p /= p.sum()
        while (np.arange(lo, hi + 1) * p).sum() >= 0:
            p[0] += 0.2
            p /= p.sum()
        a = [ctypes.c_float() for _ in range(3)]
        b = [ctypes.c_float() for _ in range(3)]
        gz.lib().gz_karlin_block(gz.ptr(p), lo, hi, *[ctypes.byref(x) for x in a])
        pc = p.copy()
        r.gzr_karlin_blk(oracle_py.ptr(pc), lo, hi, *[ctypes.byref(x) for x in b])
        assert [x.value for x in a] == [x.value for x in b]


@REF
def test_search_space_matches_reference():
    r = oracle_py.ref()
    for q in [1, 10, 50, 150, 300, 1000, 35000]:
        for dl, dn in [(1000, 3), (10**6, 3000), (330 * 10**6, 10**6), (1_300_000_000, 4 * 10**6)]:
            assert gz.lib().gz_search_space(q, dl, dn) == r.gzr_search_space(q, dl, dn)


def _ref_encode():
    t = np.zeros(256, np.uint8)
    oracle_py.ref().gzr_protein_encode_table(oracle_py.ptr(t))
    return t


def _ref_query(seq_text, translated, filter_on, table, matrix, nl):
    """Reference framing: ProteinQuery::SetEncodedSequence (protein_query.cpp:40-49),
    TranslatedDnaQuery::SetEncodedSequence (translated_dna_query.cpp:49-79)."""
    r = oracle_py.ref()

    def seg(s):
        if not filter_on:
            return 0, s
        buf = ctypes.create_string_buffer(len(s) + 1)
        st = r.gzr_seg(s, buf)
        return st, buf.raw[:len(s)]

    if not translated:
        out = np.zeros(len(seq_text) + 2, np.uint8)
        out[0] = 25
        st, masked = seg(seq_text)
        if st == 0:
            out[1:-1] = table[np.frombuffer(masked, np.uint8)] if masked else []
            out[-1] = 25
        return out
    n = len(seq_text)
    frames = ctypes.create_string_buffer(2 * n + 16)
    lens = np.zeros(6, np.uint32)
    r.gzr_translate(seq_text, frames, oracle_py.ptr(lens))
    base = n // 3
    out = np.zeros(7 + 6 * base, np.uint8)
    out[0] = 25
    off, src = 1, 0
    for i in range(6):
        s = frames.raw[src:src + int(lens[i])]
        src += int(lens[i])
        st, masked = seg(s)
        if st == 0:
            out[off:off + len(s)] = table[np.frombuffer(masked, np.uint8)] if s else []
        else:
            out[off:off + len(s)] = 24
        off += len(s)
        if len(s) < base:
            out[off] = 25
            off += 1
        out[off] = 25
        off += 1
    return out


@REF
@pytest.mark.parametrize("translated", [False, True])
@pytest.mark.parametrize("filter_on", [True, False])
def test_queries_match_reference_composition(translated, filter_on):
    rng = np.random.default_rng(14 + translated * 2 + filter_on)
    r = oracle_py.ref()
    m, nl = gz.blosum62()
    table = _ref_encode()
    texts = []
    for i in range(40):
        if translated:
            n = int(rng.integers(0, 200)) if i % 4 else 150
            texts.append("".join(rng.choice(list("ACGTACGTACGTN"), n)).encode())
        else:
            n = int(rng.integers(1, 300)) if i % 6 else int(rng.integers(1, 12))
            texts.append(_rand_protein_text(rng, n, low_complexity=bool(i % 3 == 0)).encode())
    names = [f"q{i}" for i in range(len(texts))]
    offs = np.zeros(len(texts) + 1, np.uint64)
    np.cumsum([len(t) for t in texts], out=offs[1:])
    q = gz.Queries(names, b"".join(texts), offs, translated=translated, filter=filter_on)
    v = q.view()
    li, ki, hi = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    r.gzr_ideal_karlin(oracle_py.ptr(m), nl, ctypes.byref(li), ctypes.byref(ki), ctypes.byref(hi))
    assert v["number_queries"] == len(texts)
    assert v["names"] == [n.encode() for n in names]
    for i, t in enumerate(texts):
        want = _ref_query(t, translated, filter_on, table, m, nl)
        got = v["codes"][v["offsets"][i]:v["offsets"][i + 1]]
        assert np.array_equal(got, want), (i, t)
        lb, kb, hb = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ok = r.gzr_query_karlin(oracle_py.ptr(want), len(want), oracle_py.ptr(m), nl,
                                ctypes.byref(lb), ctypes.byref(kb), ctypes.byref(hb))
        if not ok:
            lb, kb = li, ki
        assert (v["lambda"][i], v["K"][i]) == (np.float32(lb.value), np.float32(kb.value))
        if translated:
            assert v["dna_lengths"][i] == len(t)


def _ref_index(seq, clustering, L=10, threshold=39):
    r = oracle_py.ref()
    m, nl = gz.blosum62()
    h = r.gzr_build_index(oracle_py.ptr(seq), len(seq), int(clustering), L, threshold,
                          oracle_py.ptr(m), nl, b"A KR EDNQ C G H ILVM FYW P ST")
    sizes = np.zeros(7, np.uint64)
    r.gzr_index_sizes(h, oracle_py.ptr(sizes))
    arrays = [np.zeros(max(int(s), 1), np.uint32) for s in sizes]
    ptrs = (ctypes.c_void_p * 7)(*[a.ctypes.data for a in arrays])
    r.gzr_index_copy(h, ptrs)
    r.gzr_index_free(h)
    keys = ["hashes", "rep_offsets", "rep_positions", "rep_member_offsets", "member_positions",
            "value_offsets", "value_positions"]
    return {k: a[:int(s)] for k, a, s in zip(keys, arrays, sizes)}


@REF
@pytest.mark.parametrize("clustering", [True, False])
def test_database_index_matches_reference(clustering):
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261202, 300, 20, homologs=6)
    # a few exact and near duplicates so that clusters have members
    db = gz.Database(dn, dr, do, clustering=clustering, chunk_size=40000)
    info = db.info()
    assert info.number_chunks >= 2
    assert info.total_sequences == len(dn)
    members = 0
    for c in range(info.number_chunks):
        ch = db.chunk(c)
        want = _ref_index(ch["sequence"], clustering)
        for k, w in want.items():
            if not clustering and k in ("rep_positions", "member_positions"):
                assert ch[k].size == 0
                continue
            if not clustering and k in ("rep_offsets", "rep_member_offsets"):
                continue
            assert np.array_equal(ch[k], w), (c, k, ch[k][:10], w[:10])
        members += ch["member_positions"].size
    if clustering:
        assert members > 0


def test_chunking_follows_cumulative_residue_rule():
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(5, 200, 5)
    lens = np.diff(do.astype(np.int64))
    db = gz.Database(dn, dr, do, chunk_size=20000)
    info = db.info()
    seen = 0
    for c in range(info.number_chunks):
        ch = db.chunk(c)
        k = ch["number_sequences"]
        body = ch["sequence"]
        # leading delimiter, each sequence followed by one delimiter
        assert body[0] == 25 and len(body) == 1 + int(lens[seen:seen + k].sum()) + k
        assert ch["names"][0] == dn[seen].encode()
        seen += k
    assert seen == len(dn)


def test_forward_features_are_reverse_features_swapped():
    """The forward side of a hit (positions q, d) sees the reverse side's
    (q - 1, d - 1) windows in the other order: along the extension, then away
    from it.  The aligner relies on this to predict both sides from one
    feature pass (the network reads the forward row with its halves
    swapped)."""
    from helpers import related_pair_set
    rng = np.random.default_rng(13)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 3000, min_len=5, max_len=90)
    rev = gz.cost_features(q, d, np.ascontiguousarray(qpos - 1, np.uint32),
                           np.ascontiguousarray(dpos - 1, np.uint32), 1)
    fwd = gz.cost_features(q, d, np.ascontiguousarray(qpos, np.uint32),
                           np.ascontiguousarray(dpos, np.uint32), 0)
    assert np.array_equal(fwd, np.concatenate([rev[:, 64:], rev[:, :64]], axis=1))
