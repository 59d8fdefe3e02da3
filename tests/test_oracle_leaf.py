"""The oracle's leaf algorithms, pinned.

1. Known answers of the reference's own unit tests (test/*.cpp, test/*.cu).
2. Bit-exact agreement with the reference's sources compiled into oracle/_ref
   on randomised inputs (marked `ref`; skipped where oracle/_ref is absent).
3. Committed golden fixtures written from oracle/_ref (tests/golden/), so the
   pinning also holds where the reference tree is not available.
"""
import ctypes
import os

import numpy as np
import pytest

import gz_native
import oracle_py
from helpers import DELIM, concat, random_protein, related_pair_set

INT_MAX = 2**31 - 1
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# DNA test alphabet of the reference tests: A C G T = 0..3, N = 4, delimiter 5,
# ScoreMatrix("test_matrix", 4, 2, -1).
DNA = {"A": 0, "C": 1, "G": 2, "T": 3}
DNA_DELIM = 5
TEST_MATRIX = np.array([[2 if i == j else -1 for j in range(4)] for i in range(4)],
                       np.int32).ravel()


def dna(s):
    return np.array([DNA_DELIM] + [DNA[c] for c in s] + [DNA_DELIM], np.uint8)


def addr(a, off=0):
    return ctypes.c_void_p(a.ctypes.data + int(off))


# ---- 1. reference unit-test known answers ----------------------------------
def test_ungapped_known_answers():
    """test/ungapped_extender_test.cpp:18-116."""
    O = oracle_py.oracle()
    s0, s1 = dna("AAAAAAC"), dna("AACAAAG")
    p0, p1 = ctypes.c_int(), ctypes.c_int()
    score = O.gzo_ungapped_one_side_positions(addr(s0, 1), addr(s1, 1), DNA_DELIM, 0,
                                              oracle_py.ptr(TEST_MATRIX), 4, INT_MAX,
                                              ctypes.byref(p0), ctypes.byref(p1))
    assert (score, p0.value, p1.value) == (9, 5, 5)
    rscore = O.gzo_ungapped_one_side_positions(addr(s0, s0.size - 3), addr(s1, s1.size - 3),
                                               DNA_DELIM, 1, oracle_py.ptr(TEST_MATRIX), 4,
                                               INT_MAX, ctypes.byref(p0), ctypes.byref(p1))
    assert (rscore, p0.value, p1.value) == (9, -5, -5)
    score = O.gzo_ungapped_one_side_positions(addr(s0, 1), addr(s1, 1), DNA_DELIM, 0,
                                              oracle_py.ptr(TEST_MATRIX), 4, 1,
                                              ctypes.byref(p0), ctypes.byref(p1))
    assert (score, p0.value, p1.value) == (4, 1, 1)
    # score-only variant agrees
    assert O.gzo_ungapped_one_side(addr(s0, 1), addr(s1, 1), DNA_DELIM, 0,
                                   oracle_py.ptr(TEST_MATRIX), 4, INT_MAX) == 9


def test_gapped_known_answer():
    """test/gapped_extender_test.cpp:24-93: score 7 ending at (4, 3), ops S S G1 S S."""
    s0, s1 = dna("AGTCAC"), dna("AGCAG")
    s, p0, p1, edits, _ = oracle_py.gapped_one_side(
        "oracle", s0, 1, s1.size - 1, s1, 1, DNA_DELIM, False, TEST_MATRIX, 4, 0, 1, INT_MAX,
        traceback=True)
    assert (s, p0, p1) == (7, 4, 3)
    assert edits == [1, 1, 2, 1, 1]  # kSubstitution=1, kGapInSeq1=2 (edit_blocks.h:14-18)
    # score-only path gives the same end point
    s2, q2, d2, _, cells = oracle_py.gapped_one_side(
        "oracle", s0, 1, s1.size - 1, s1, 1, DNA_DELIM, False, TEST_MATRIX, 4, 0, 1, INT_MAX)
    assert (s2, q2, d2) == (7, 4, 3) and cells > 0


def test_gapped_gpu_test_answers():
    """test/gapped_extender_gpu_test.cu:40-253 (seed at position 1; positions
    reported there as seed + best offset)."""
    s0, s1 = dna("AGCAC"), dna("AGCAG")
    s, p0, p1, _, _ = oracle_py.gapped_one_side("oracle", s0, 1, s0.size, s1, 1, DNA_DELIM,
                                                False, TEST_MATRIX, 4, 0, 1, 0)
    assert (s, 1 + p0, 1 + p1) == (8, 4, 4)
    s0 = dna("AGTCAC")
    s, p0, p1, _, _ = oracle_py.gapped_one_side("oracle", s0, 1, s0.size, s1, 1, DNA_DELIM,
                                                False, TEST_MATRIX, 4, 0, 1, INT_MAX)
    assert (s, 1 + p0, 1 + p1) == (7, 5, 4)


def test_chain_filter_known_answers():
    """test/chain_filter_test.cpp:11-121."""
    O = oracle_py.oracle()

    def run(q, d, hits):
        s0, s1 = dna(q), dna(d)
        h = np.array(hits, np.uint32).ravel()
        return O.gzo_chain_filter(oracle_py.ptr(s0), s0.size, 1, oracle_py.ptr(s1), DNA_DELIM,
                                  oracle_py.ptr(TEST_MATRIX), 4, oracle_py.ptr(h), len(hits))

    assert run("AAAAAAC", "AAACAAG", [(1, 1), (3, 3)]) == 1
    assert run("AAAAAAC", "AACAAAG", [(1, 1), (2, 2)]) == 1
    assert run("AAAAAAC", "AACAAAG", [(1, 1), (4, 4)]) == 2


def test_distance_known_answers():
    """test/distance_calculator_test.cpp:56-117 (reduced alphabet
    "A KR EDNQ C G H ILVM FYW P ST")."""
    O = oracle_py.oracle()
    rmap, _ = gz_native.reduced_map()
    enc = gz_native.encode_protein
    a, b = enc("GMEHQS"), enc("GMEHQS")
    assert O.gzo_distance(oracle_py.ptr(a), 3, oracle_py.ptr(b), 3, 6, oracle_py.ptr(rmap)) == 0
    a, b = enc("IMEHQS"), enc("LMEHQA")
    assert O.gzo_distance(oracle_py.ptr(a), 3, oracle_py.ptr(b), 3, 6, oracle_py.ptr(rmap)) == 1
    a, b = enc("IMEHQSIMEHQS"), enc("LMEHQALMEHQA")
    got = [O.gzo_distance(oracle_py.ptr(a), c, oracle_py.ptr(b), c, 4, oracle_py.ptr(rmap))
           for c in (2, 6, 10)]
    assert got == [0, 1, 1]


def test_gapped_textbook_without_xdrop():
    """With cutoff = INT_MAX the one-sided extension is the textbook Gotoh
    affine DP anchored at the seed: best score over all prefixes pairs,
    checked by brute force on small random inputs."""
    rng = np.random.default_rng(7)
    m, nl = gz_native.blosum62()
    M = m.reshape(25, 25)
    for _ in range(60):
        a = random_protein(rng, int(rng.integers(1, 14)))
        b = random_protein(rng, int(rng.integers(1, 14)))
        s0, s1 = concat([a])[0], concat([b])[0]
        go, ge = 11, 1
        NEG = -10**9
        n, k = len(a), len(b)
        H = np.full((k + 1, n + 1), NEG, np.int64)
        E = np.full((k + 1, n + 1), NEG, np.int64)
        F = np.full((k + 1, n + 1), NEG, np.int64)
        H[0, 0] = 0
        for j in range(1, n + 1):
            E[0, j] = -(go + ge * j)
            H[0, j] = E[0, j]
        for i in range(1, k + 1):
            F[i, 0] = -(go + ge * i)
            H[i, 0] = F[i, 0]
            for j in range(1, n + 1):
                E[i, j] = max(E[i, j - 1] - ge, H[i, j - 1] - go - ge)
                F[i, j] = max(F[i - 1, j] - ge, H[i - 1, j] - go - ge)
                H[i, j] = max(H[i - 1, j - 1] + M[b[i - 1], a[j - 1]], E[i, j], F[i, j])
        best = max(0, int(H[1:, 1:].max()))
        s, _, _, _, _ = oracle_py.gapped_one_side("oracle", s0, 1, s0.size, s1, 1, DELIM, False,
                                                  m, nl, go, ge, INT_MAX)
        assert s == best


# ---- 2. bit-exact against the reference sources (oracle/_ref) ---------------
@pytest.mark.ref
def test_blosum62_and_alphabet_match_reference():
    R = oracle_py.ref()
    m, nl = gz_native.blosum62()
    mat = np.zeros(625, np.int32)
    n = ctypes.c_uint32()
    R.gzr_read_matrix(gz_native.blosum62_text(), oracle_py.ptr(mat), ctypes.byref(n))
    assert n.value == nl and np.array_equal(mat, m)
    t = np.zeros(256, np.uint8)
    R.gzr_protein_encode_table(oracle_py.ptr(t))
    mine = gz_native.encode_table()
    assert np.array_equal(t[:128], mine[:128])
    rm = np.zeros(25, np.uint8)
    mx = ctypes.c_uint32()
    R.gzr_reduced_map(b"A KR EDNQ C G H ILVM FYW P ST", oracle_py.ptr(rm), ctypes.byref(mx))
    mine_map, mine_mx = gz_native.reduced_map()
    assert np.array_equal(rm, mine_map) and mx.value == mine_mx


@pytest.mark.ref
def test_extenders_match_reference_randomised():
    rng = np.random.default_rng(11)
    m, nl = gz_native.blosum62()
    O, R = oracle_py.oracle(), oracle_py.ref()
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 300, max_len=200)
    for i in range(len(qpos)):
        for rev, (a, b) in ((1, (qpos[i] - 1, dpos[i] - 1)), (0, (qpos[i], dpos[i]))):
            for cutoff in (int(rng.integers(1, 40)), INT_MAX):
                o = O.gzo_ungapped_one_side(addr(q, a), addr(d, b), DELIM, rev,
                                            oracle_py.ptr(m), nl, cutoff)
                r = R.gzr_ungapped_one_side(addr(q, a), addr(d, b), DELIM, rev,
                                            oracle_py.ptr(m), nl, cutoff)
                assert o == r
            # gapped, score-only and traceback
            seq0_len = a if rev else q.size - a
            cutoff = int(rng.choice([1, 5, 15, 26, 40, 52]))
            go, ge = (11, 1) if rng.random() < 0.8 else (int(rng.integers(0, 12)),
                                                        int(rng.integers(1, 4)))
            for tb in (False, True):
                o = oracle_py.gapped_one_side("oracle", q, a, seq0_len, d, b, DELIM, rev, m, nl,
                                              go, ge, cutoff, tb)
                r = oracle_py.gapped_one_side("ref", q, a, seq0_len, d, b, DELIM, rev, m, nl,
                                              go, ge, cutoff, tb)
                assert o[:4] == r[:4], (i, rev, tb)


@pytest.mark.ref
def test_distance_and_hash_match_reference_randomised():
    rng = np.random.default_rng(5)
    O, R = oracle_py.oracle(), oracle_py.ref()
    rmap, max_regular = gz_native.reduced_map()
    a = random_protein(rng, 5000, ambiguous=0.02)
    b = random_protein(rng, 5000, ambiguous=0.02)
    for _ in range(2000):
        i, j = int(rng.integers(5, 4990)), int(rng.integers(5, 4990))
        L = int(rng.choice([4, 6, 10]))
        assert O.gzo_distance(oracle_py.ptr(a), i, oracle_py.ptr(b), j, L, oracle_py.ptr(rmap)) == \
            R.gzr_distance(oracle_py.ptr(a), i, oracle_py.ptr(b), j, L, oracle_py.ptr(rmap))
    # hash function state as the db builder derives it
    m, nl = gz_native.blosum62()
    M = m.reshape(25, 25)
    scores = np.zeros(max_regular + 1, np.int32)
    for c in range(25):
        r = rmap[c]
        if r <= max_regular:
            scores[r] = max(scores[r], M[c, c])
    shifted_map = (rmap.astype(np.int32) + 1).astype(np.uint8)
    shifted_scores = np.concatenate([[0], scores]).astype(np.int32)
    seq = np.concatenate([random_protein(rng, 3000, ambiguous=0.03), [DELIM] * 8]).astype(np.uint8)
    for i in range(3000):
        h1, l1 = ctypes.c_uint32(), ctypes.c_uint32()
        h2, l2 = ctypes.c_uint32(), ctypes.c_uint32()
        r1 = O.gzo_hash(addr(seq, i), oracle_py.ptr(shifted_map), 25,
                        oracle_py.ptr(shifted_scores), max_regular + 1, 39, 8, 4,
                        ctypes.byref(h1), ctypes.byref(l1))
        r2 = R.gzr_hash(addr(seq, i), oracle_py.ptr(rmap), max_regular, 39, 8,
                        oracle_py.ptr(scores), ctypes.byref(h2), ctypes.byref(l2))
        assert r1 == r2
        if r1 == 0:
            assert (h1.value, l1.value) == (h2.value, l2.value)


# ---- 3. committed golden fixtures --------------------------------------------
def test_oracle_matches_golden_gapped_fixture():
    path = os.path.join(GOLDEN, "gapped_extender.npz")
    if not os.path.exists(path):
        pytest.skip("golden fixture not generated")
    g = np.load(path)
    m, nl = gz_native.blosum62()
    q, d = g["query"], g["database"]
    for k in range(g["qstart"].size):
        o = oracle_py.gapped_one_side("oracle", q, int(g["qstart"][k]), int(g["seq0_len"][k]), d,
                                      int(g["dstart"][k]), DELIM, int(g["reverse"][k]), m, nl,
                                      int(g["gap_open"][k]), int(g["gap_ext"][k]),
                                      int(g["cutoff"][k]))
        assert (o[0], o[1], o[2]) == (g["score"][k], g["p0"][k], g["p1"][k])
