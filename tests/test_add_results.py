"""AlignerCommon::AddResults (src/aligner_common.h:147-274,
src/aligner_common.cpp:12-68), the merge of one database chunk's presearched
results into a query's kept list, checked in isolation on the CPU.

The expected lists below were worked out by hand from the reference source,
one case per rule: the cap of -b kept results, the cap of -v per subject of
the current chunk, AlignmentComp's order (score, then query length, then
subject length, then the smaller subject start), the stable sort of the new
results, the merge with the earlier chunks' results (a full tie takes the new
result first), and the containment check, including the reference's quirk:
it stores each kept result's *query* start in the per-subject list but
searches that list with *subject* positions, so a contained result is only
caught when the kept one's query start lies in [start.d, end.d) of the new
one.  The oracle (gzo_add_results) and the product (results.cpp, through
libhostpipe's gzh_add_results) must both give every listed answer; then they
are compared with each other and with the merge's invariants on random
inputs full of ties.
"""
import ctypes
from ctypes import POINTER, Structure, c_int, c_uint32, c_void_p

import numpy as np
import pytest

import oracle_py


class Pre(Structure):
    _fields_ = [("chunk", c_int), ("subject", c_uint32), ("score", c_int),
                ("hit_q", c_uint32), ("hit_d", c_uint32), ("start_q", c_uint32),
                ("start_d", c_uint32), ("end_q", c_uint32), ("end_d", c_uint32),
                ("hit_count", c_uint32)]


ARGS = [c_void_p, c_uint32, c_int, c_uint32, c_uint32, c_void_p, c_uint32, c_void_p,
        POINTER(c_uint32)]
# 20 subjects of 99 residues: subject i holds chunk positions 1 + 100 i ..
OFFSETS = np.array([1 + 100 * i for i in range(21)], np.uint32)


def lib(kind):
    L = oracle_py.oracle() if kind == "oracle" else oracle_py.hostpipe()
    f = L.gzo_add_results if kind == "oracle" else L.gzh_add_results
    f.restype = c_int
    f.argtypes = ARGS
    return f


def R(score, start, end, hit=(0, 0), chunk=-1, subject=0, tag=0):
    return (chunk, subject, score, hit[0], hit[1], start[0], start[1], end[0], end[1], tag)


def add_results(kind, added, results=(), chunk_id=1, b=10, v=1, offsets=OFFSETS):
    f = lib(kind)
    A = (Pre * max(1, len(added)))(*[Pre(*x) for x in added])
    cap = b + len(results) + len(added) + 1
    Rs = (Pre * cap)(*[Pre(*x) for x in results])
    n = c_uint32(len(results))
    assert f(offsets.ctypes.data, len(offsets) - 1, chunk_id, b, v, ctypes.addressof(A),
             len(added), ctypes.addressof(Rs), ctypes.byref(n)) == 0
    return [tuple(getattr(Rs[i], k) for k, _ in Pre._fields_) for i in range(n.value)]


def tags(out):
    return [r[9] for r in out]


KINDS = ["oracle", "product"]


@pytest.mark.parametrize("kind", KINDS)
def test_order_and_result_cap(kind):
    added = [
        R(50, (10, 105), (40, 135), tag=0),   # q 30, s 30, start 105
        R(60, (10, 205), (20, 215), tag=1),
        R(50, (10, 305), (45, 340), tag=2),   # longer on the query
        R(50, (10, 405), (40, 440), tag=3),   # q 30, longer on the subject
        R(50, (10, 5), (40, 35), tag=4),      # q 30, s 30, smaller start
        R(10, (10, 505), (20, 515), tag=5),
        R(20, (10, 605), (20, 615), tag=6),
        R(30, (10, 705), (20, 715), tag=7),
        R(40, (10, 805), (20, 815), tag=8),
        R(45, (10, 905), (20, 915), tag=9),
        R(5, (10, 1005), (20, 1015), tag=10),
        R(1, (10, 1105), (20, 1115), tag=11),
    ]
    out = add_results(kind, added, b=10)
    assert tags(out) == [1, 2, 3, 4, 0, 9, 8, 7, 6, 5]
    # chunk id and subject (DatabaseChunk::GetId) are filled in
    assert [(r[0], r[1]) for r in out[:5]] == [(1, 2), (1, 3), (1, 4), (1, 0), (1, 1)]


@pytest.mark.parametrize("kind", KINDS)
def test_per_subject_cap(kind):
    added = [R(30, (10, 110), (20, 120), tag=0), R(20, (30, 130), (40, 140), tag=1),
             R(10, (50, 150), (60, 160), tag=2)]
    assert tags(add_results(kind, added, v=1)) == [0]
    assert tags(add_results(kind, added, v=2)) == [0, 1]
    assert tags(add_results(kind, added, v=3)) == [0, 1, 2]


@pytest.mark.parametrize("kind", KINDS)
def test_dropped_results_do_not_count_against_the_cap(kind):
    added = [R(30, (10, 110), (20, 120), tag=0), R(20, (30, 130), (40, 140), tag=1),
             R(10, (10, 210), (20, 220), tag=2)]
    assert tags(add_results(kind, added, b=2, v=1)) == [0, 2]


@pytest.mark.parametrize("kind", KINDS)
def test_containment_as_the_reference_checks_it(kind):
    # contained, but the kept query start (10) is not in [112, 125): kept
    c1 = [R(20, (10, 110), (30, 130), tag=0), R(15, (12, 112), (25, 125), tag=1)]
    assert tags(add_results(kind, c1, v=2)) == [0, 1]
    # contained and the kept query start (120) lies in [110, 130): dropped
    c2 = [R(20, (120, 105), (150, 140), tag=0), R(15, (125, 110), (140, 130), tag=1)]
    assert tags(add_results(kind, c2, v=2)) == [0]
    # same start (IsContained), but 120 is not in [105, 115): kept
    c3 = [R(20, (120, 105), (150, 140), tag=0), R(15, (120, 105), (130, 115), tag=1)]
    assert tags(add_results(kind, c3, v=2)) == [0, 1]
    # containment with equality on the query end and the subject start: dropped
    c5 = [R(20, (120, 105), (150, 140), tag=0), R(15, (121, 105), (150, 139), tag=1)]
    assert tags(add_results(kind, c5, v=2)) == [0]
    # not contained (the new one ends later on the subject): kept
    c6 = [R(20, (120, 105), (150, 140), tag=0), R(15, (121, 110), (150, 141), tag=1)]
    assert tags(add_results(kind, c6, v=2)) == [0, 1]


@pytest.mark.parametrize("kind", KINDS)
def test_merge_with_earlier_chunks(kind):
    prev = [R(30, (1, 5), (9, 13), chunk=0, subject=1, tag=100),
            R(10, (1, 5), (9, 13), chunk=0, subject=2, tag=101)]
    added = [R(20, (10, 110), (20, 120), tag=0), R(5, (10, 210), (20, 220), tag=1)]
    out = add_results(kind, added, prev, chunk_id=1, b=3, v=1)
    assert tags(out) == [100, 0, 101]
    # an earlier chunk's result on "subject 1" is another subject: both kept
    assert [(r[0], r[1]) for r in out] == [(0, 1), (1, 1), (0, 2)]


@pytest.mark.parametrize("kind", KINDS)
def test_full_tie_takes_the_new_result_first(kind):
    prev = [R(10, (5, 105), (15, 115), chunk=0, subject=1, tag=100)]
    added = [R(10, (5, 105), (15, 115), tag=0)]
    assert tags(add_results(kind, added, prev, b=10, v=1)) == [0, 100]


@pytest.mark.parametrize("kind", KINDS)
def test_new_results_sort_stably(kind):
    a1 = R(10, (5, 105), (15, 115), hit=(7, 107), tag=1)
    a2 = R(10, (5, 105), (15, 115), hit=(8, 108), tag=2)
    assert tags(add_results(kind, [a1, a2], v=2)) == [1, 2]
    assert tags(add_results(kind, [a2, a1], v=2)) == [2, 1]


def comp(a, b):
    """AlignmentComp's sign for two result tuples (for the invariants)."""
    if a[2] != b[2]:
        return 1 if a[2] > b[2] else -1
    aq, bq = (a[7] - a[5]) & 0xffffffff, (b[7] - b[5]) & 0xffffffff
    if aq != bq:
        return 1 if aq > bq else -1
    ad, bd = (a[8] - a[6]) & 0xffffffff, (b[8] - b[6]) & 0xffffffff
    if ad != bd:
        return 1 if ad > bd else -1
    return 1 if a[6] < b[6] else -1


def random_case(rng, chunk_id):
    n = int(rng.integers(0, 40))
    added = []
    for t in range(n):
        s = int(rng.integers(1, 19))
        sd = 1 + 100 * s + int(rng.integers(0, 30))
        sq = int(rng.integers(0, 200))
        L = int(rng.integers(1, 4)) * 5
        dq, dd = L + int(rng.integers(0, 2)), L + int(rng.integers(0, 2))
        added.append(R(int(rng.integers(5, 9)), (sq, sd), (sq + dq, sd + dd),
                       hit=(sq + 1, sd + 1), tag=t))
    return added


@pytest.mark.parametrize("seed", range(4))
def test_oracle_and_product_agree_on_random_merges(seed):
    rng = np.random.default_rng(9000 + seed)
    for case in range(300):
        b = int(rng.integers(1, 12))
        v = int(rng.integers(1, 4))
        kept = {"oracle": [], "product": []}
        for chunk_id in range(3):
            added = random_case(rng, chunk_id)
            outs = {k: add_results(k, added, kept[k], chunk_id=chunk_id, b=b, v=v)
                    for k in KINDS}
            assert outs["oracle"] == outs["product"], (seed, case, chunk_id)
            out = outs["oracle"]
            prev = kept["oracle"]
            assert len(out) <= b
            # at most v per subject of this chunk
            subj = [r[1] for r in out if r[0] == chunk_id]
            assert all(subj.count(s) <= v for s in set(subj))
            # sorted: no result comes after one it beats
            assert all(comp(out[i + 1], out[i]) < 0 for i in range(len(out) - 1))
            # earlier results keep their relative order and come first at ties
            olds = [r for r in out if r[0] != chunk_id]
            assert olds == [r for r in prev if r in olds]
            # as many as the caps allow are kept: the list is short of b only
            # when every candidate was either kept or dropped by a cap
            if len(out) < b:
                assert len(olds) == len(prev)
            kept = {k: outs[k] for k in KINDS}
