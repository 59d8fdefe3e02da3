"""HIP extender kernels vs the CPU oracle, through the C ABI.

Every comparison is bit-exact: scores, best positions, flags and distances
are integers.  Inputs are seeded: homologous pairs (with indels) so that
extensions are long and gapped, plus unrelated pairs for early X-drop stops.
"""
import ctypes
import sys

import numpy as np
import pytest

import gz_native as gz
import oracle_py
from helpers import DELIM, concat, random_protein, related_pair_set

pytestmark = pytest.mark.gpu
INT_MAX = 2**31 - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    gz.check(gz.lib().gz_set_device(0))
    m, nl = gz.blosum62()
    return {"matrix": m, "nl": nl, "d_matrix": gz.DeviceBuffer.from_array(m)}


def test_pack_round_trip(dev):
    rng = np.random.default_rng(1)
    for n in (0, 1, 11, 12, 13, 47, 48, 49, 1000, 12345):
        say("pack n", n)
        codes = rng.integers(0, 26, n).astype(np.uint8)
        _, packed = gz.pack_on_device(codes)
        words = gz.lib().gz_packed_words(n)
        host = packed.to_array(np.uint64, words)
        got = np.array([gz.lib().gz_packed_code_at(gz.ptr(host), i) for i in range(n)], np.uint8)
        assert np.array_equal(got, codes)
        # padding reads as delimiter in front and behind
        if n:
            assert gz.lib().gz_packed_code_at(gz.ptr(host), n) == DELIM


def test_distance_kernel(dev):
    rng = np.random.default_rng(2)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 400, min_len=30)
    rmap, _ = gz.reduced_map()
    # centres at least L/2 inside their sequences
    qc, dc = [], []
    for i in range(len(qpos)):
        for _ in range(8):
            a = int(rng.integers(qo[i] + 5, qo[i + 1] - 6)) if qo[i + 1] - qo[i] > 12 else None
            b = int(rng.integers(do[i] + 5, do[i + 1] - 6)) if do[i + 1] - do[i] > 12 else None
            if a is None or b is None:
                continue
            qc.append(a)
            dc.append(b if rng.random() < 0.5 else min(max(a - qo[i] + do[i], do[i] + 5),
                                                       do[i + 1] - 6))
    qc, dc = np.array(qc, np.uint32), np.array(dc, np.uint32)
    n = qc.size
    say("distance pairs", n)
    _, qp = gz.pack_on_device(q)
    _, dp = gz.pack_on_device(d)
    d_map = gz.DeviceBuffer.from_array(rmap)
    h = ctypes.c_void_p()
    L = gz.lib()
    gz.check(L.gz_distance_calculator_create(ctypes.byref(h)))
    gz.check(L.gz_distance_calculator_set_queries(h, qp.ptr))
    gz.check(L.gz_distance_calculator_set_database(h, dp.ptr))
    gz.check(L.gz_distance_calculator_set_reduced_code_map(h, d_map.ptr))
    gz.check(L.gz_distance_calculator_set_subsequence_length(h, 10))
    d_qc, d_dc, d_out = gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n)
    out = np.zeros(n, np.int8)
    s = gz.Stream()
    gz.check(L.gz_distance_calculator_calculate_distances_async(
        h, n, gz.ptr(qc), gz.ptr(dc), gz.ptr(out), d_qc.ptr, d_dc.ptr, d_out.ptr, s.handle))
    s.sync()
    O = oracle_py.oracle()
    want = np.array([O.gzo_distance(gz.ptr(q), int(a), gz.ptr(d), int(b), 10, gz.ptr(rmap))
                     for a, b in zip(qc, dc)], np.int8)
    assert np.array_equal(out, want)
    L.gz_distance_calculator_destroy(h)


def _ungapped_case(dev, rng, nseq, related=True):
    if related:
        q, qo, d, do, qpos, dpos = related_pair_set(rng, nseq, min_len=10, max_len=400)
    else:
        qs = [random_protein(rng, int(rng.integers(5, 300)), 0.01) for _ in range(nseq)]
        ds = [random_protein(rng, int(rng.integers(5, 300)), 0.01) for _ in range(nseq)]
        q, qo = concat(qs)
        d, do = concat(ds)
        qpos = np.array([rng.integers(qo[i], qo[i + 1] - 1) for i in range(nseq)], np.uint32)
        dpos = np.array([rng.integers(do[i], do[i + 1] - 1) for i in range(nseq)], np.uint32)
    qids = np.arange(nseq, dtype=np.uint32)
    cut = rng.integers(5, 30, nseq).astype(np.int32)
    trig = rng.integers(10, 60, nseq).astype(np.int32)
    return q, d, qids, qpos, dpos, cut, trig


@pytest.mark.parametrize("related", [True, False])
def test_ungapped_with_trigger_kernel(dev, related):
    rng = np.random.default_rng(3 if related else 4)
    q, d, qids, qpos, dpos, cut, trig = _ungapped_case(dev, rng, 3000, related)
    n = qids.size
    say("ungapped seeds", n, "related", related)
    _, qp = gz.pack_on_device(q)
    _, dp = gz.pack_on_device(d)
    d_cut, d_trig = gz.DeviceBuffer.from_array(cut), gz.DeviceBuffer.from_array(trig)
    L = gz.lib()
    h = ctypes.c_void_p()
    gz.check(L.gz_ungapped_extender_create(ctypes.byref(h)))
    gz.check(L.gz_ungapped_extender_set_queries(h, DELIM, qp.ptr, d_cut.ptr, d_trig.ptr))
    gz.check(L.gz_ungapped_extender_set_database(h, dp.ptr))
    gz.check(L.gz_ungapped_extender_set_score_matrix(h, dev["d_matrix"].ptr, dev["nl"]))
    bufs = [gz.DeviceBuffer(n * 4) for _ in range(3)]
    d_flags = gz.DeviceBuffer(n)
    flags = np.zeros(n, np.int8)
    s = gz.Stream()
    gz.check(L.gz_ungapped_extender_extend_with_trigger_async(
        h, n, gz.ptr(qids), gz.ptr(qpos), gz.ptr(dpos), gz.ptr(flags), bufs[0].ptr, bufs[1].ptr,
        bufs[2].ptr, d_flags.ptr, None, s.handle))
    s.sync()
    O = oracle_py.oracle()
    m = dev["matrix"]
    want = np.array([O.gzo_ungapped_with_trigger(gz.ptr(q), int(a), gz.ptr(d), int(b), DELIM,
                                                 gz.ptr(m), 25, int(c), int(t)) > int(t)
                     for a, b, c, t in zip(qpos, dpos, cut, trig)], np.int8)
    assert np.array_equal(flags, want)
    if related:
        assert want.sum() > n // 10  # the test exercises passing seeds
    L.gz_ungapped_extender_destroy(h)


def test_set_query_seeds_expansion(dev):
    """Run-length seeds expand to per-seed query id / position (max-scan)."""
    rng = np.random.default_rng(5)
    L = gz.lib()
    nseed_table = 5000
    table_ids = rng.integers(0, 1000, nseed_table).astype(np.uint32)
    table_pos = rng.integers(0, 10**7, nseed_table).astype(np.uint32)
    d_tid, d_tpos = gz.DeviceBuffer.from_array(table_ids), gz.DeviceBuffer.from_array(table_pos)
    h = ctypes.c_void_p()
    gz.check(L.gz_ungapped_extender_create(ctypes.byref(h)))
    gz.check(L.gz_ungapped_extender_set_query_seed_data_list(h, d_tid.ptr, d_tpos.ptr))
    for runs in (1, 7, 1000, 70000):
        lens = rng.integers(1, 9, runs)
        starts = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32)
        size = int(lens.sum())
        ids = rng.integers(0, nseed_table, runs).astype(np.uint32)
        say("expansion runs", runs, "seeds", size)
        tb = L.gz_ungapped_extender_query_seeds_temp_bytes(size)
        d_temp = gz.DeviceBuffer(tb)
        d_ids, d_starts = gz.DeviceBuffer(runs * 4), gz.DeviceBuffer(runs * 4)
        d_qid, d_qpos, d_tmp = gz.DeviceBuffer(size * 4), gz.DeviceBuffer(size * 4), \
            gz.DeviceBuffer(size * 4)
        s = gz.Stream()
        gz.check(L.gz_ungapped_extender_set_query_seeds(
            h, runs, size, gz.ptr(ids), gz.ptr(starts), tb, d_temp.ptr, d_ids.ptr, d_starts.ptr,
            d_qid.ptr, d_qpos.ptr, d_tmp.ptr, s.handle))
        s.sync()
        run_of = np.repeat(np.arange(runs), lens)
        assert np.array_equal(d_qid.to_array(np.uint32, size), table_ids[ids[run_of]])
        assert np.array_equal(d_qpos.to_array(np.uint32, size), table_pos[ids[run_of]])
    L.gz_ungapped_extender_destroy(h)


def _gapped_gpu(dev, q, d, qpos, dpos, reverse, cutoff, gap_open=11, gap_ext=1, matrix=None,
                nl=None, delim=DELIM, narrow=None, wide_first=None, persistent=False):
    L = gz.lib()
    n = qpos.size
    _, qp = gz.pack_on_device(q)
    _, dp = gz.pack_on_device(d)
    d_m = dev["d_matrix"] if matrix is None else gz.DeviceBuffer.from_array(matrix)
    h = ctypes.c_void_p()
    gz.check(L.gz_gapped_extender_create(ctypes.byref(h)))
    gz.check(L.gz_gapped_extender_set_queries(h, delim, qp.ptr))
    gz.check(L.gz_gapped_extender_set_database(h, dp.ptr))
    gz.check(L.gz_gapped_extender_set_score_parameters(h, d_m.ptr, nl or dev["nl"], gap_open,
                                                       gap_ext, cutoff))
    if narrow is not None:
        gz.check(L.gz_gapped_extender_set_narrow_ring(h, narrow))
    if persistent:
        gz.check(L.gz_gapped_extender_set_persistent(h, 1))
    hq, hd = qpos.copy(), dpos.copy()
    hs = np.zeros(n, np.int32)
    d_q, d_d, d_s = gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n * 4)
    s = gz.Stream()
    if wide_first is None:
        gz.check(L.gz_gapped_extender_extend_one_side_score_only_async(
            h, n, int(reverse), gz.ptr(hq), gz.ptr(hd), gz.ptr(hs), d_q.ptr, d_d.ptr, d_s.ptr,
            s.handle))
        s.sync()
    else:
        d_q, d_d = gz.DeviceBuffer.from_array(hq), gz.DeviceBuffer.from_array(hd)
        gz.check(L.gz_gapped_extender_extend_one_side_score_only_binned_async(
            h, n, wide_first, int(reverse), d_q.ptr, d_d.ptr, d_s.ptr, None, s.handle))
        s.sync()
        hq, hd, hs = (d_q.to_array(np.uint32, n), d_d.to_array(np.uint32, n),
                      d_s.to_array(np.int32, n))
    L.gz_gapped_extender_destroy(h)
    return hs, hq, hd


def _gapped_oracle(q, d, qpos, dpos, reverse, cutoff, matrix, nl, gap_open=11, gap_ext=1,
                   delim=DELIM):
    out = []
    for a, b in zip(qpos, dpos):
        s, p0, p1, _, _ = oracle_py.gapped_one_side("oracle", q, int(a), q.size, d, int(b), delim,
                                                    reverse, matrix, nl, gap_open, gap_ext,
                                                    cutoff)
        out.append((s, int(a) + p0, int(b) + p1))
    return np.array(out, np.int64)


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("cutoff", [26, 52, 8])
def test_gapped_kernel_vs_oracle(dev, reverse, cutoff):
    rng = np.random.default_rng(10 + cutoff + int(reverse))
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 1500, min_len=10, max_len=600,
                                                identity=(0.25, 0.95), indel=0.04)
    if reverse:
        qpos, dpos = qpos - 1, dpos - 1  # reverse seeds start one before the hit
    say("gapped seeds", qpos.size, "reverse", reverse, "cutoff", cutoff)
    hs, hq, hd = _gapped_gpu(dev, q, d, qpos, dpos, reverse, cutoff)
    want = _gapped_oracle(q, d, qpos, dpos, reverse, cutoff, dev["matrix"], dev["nl"])
    got = np.stack([hs, hq.astype(np.int64), hd.astype(np.int64)], 1)
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, (bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("narrow,wide_first", [(16, None), (32, None), (32, 500)])
def test_gapped_kernel_ring_tiers(dev, narrow, wide_first):
    """The first LDS ring (16 or 32 columns, or the 64-column ring for the
    first wide_first seeds) only decides where a band runs: every tier gives
    the oracle's answers.  Long homologous pairs make many bands outgrow the
    narrow ring."""
    rng = np.random.default_rng(30 + narrow)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 1500, min_len=40, max_len=900,
                                                identity=(0.4, 0.98), indel=0.05)
    say("ring tiers: narrow", narrow, "wide_first", wide_first, "seeds", qpos.size)
    L = gz.lib()
    for reverse in (False, True):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        hs, hq, hd = _gapped_gpu(dev, q, d, a, b, reverse, 52, narrow=narrow,
                                 wide_first=wide_first)
        want = _gapped_oracle(q, d, a, b, reverse, 52, dev["matrix"], dev["nl"])
        got = np.stack([hs, hq.astype(np.int64), hd.astype(np.int64)], 1)
        bad = np.nonzero((got != want).any(1))[0]
        assert bad.size == 0, (bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("narrow", [16, 32])
@pytest.mark.parametrize("cutoff", [26, 8200])
def test_gapped_persistent_queue_vs_oracle(dev, narrow, cutoff):
    """The persistent work-queue launch (lanes take the next job when their
    extension ends) gives every job the oracle's score and end positions, in
    both directions, with 16- and 32-bit cells (cutoff 8200 forces 32-bit
    rings) and bands that outgrow the first ring."""
    rng = np.random.default_rng(50 + narrow + cutoff)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 2500, min_len=5, max_len=700,
                                                identity=(0.3, 0.98), indel=0.05)
    say("persistent queue: narrow", narrow, "cutoff", cutoff, "seeds", qpos.size)
    for reverse in (False, True):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        hs, hq, hd = _gapped_gpu(dev, q, d, a, b, reverse, cutoff, narrow=narrow,
                                 persistent=True)
        want = _gapped_oracle(q, d, a, b, reverse, cutoff, dev["matrix"], dev["nl"])
        got = np.stack([hs, hq.astype(np.int64), hd.astype(np.int64)], 1)
        bad = np.nonzero((got != want).any(1))[0]
        assert bad.size == 0, (reverse, bad[:5], got[bad[:5]], want[bad[:5]])


def test_gapped_kernel_ordered_jobs(dev):
    """Jobs run through an order array (the learned schedule's permutation)
    give every job the same score, end positions and DP cells as the
    unordered call, at the job's own index; an order entry past the end is
    skipped by the kernel and reported as a failure."""
    rng = np.random.default_rng(41)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 1500, min_len=40, max_len=900,
                                                identity=(0.4, 0.98), indel=0.05)
    n = qpos.size
    say("ordered gapped jobs", n)
    L = gz.lib()
    _, qp = gz.pack_on_device(q)
    _, dp = gz.pack_on_device(d)
    h = ctypes.c_void_p()
    gz.check(L.gz_gapped_extender_create(ctypes.byref(h)))
    gz.check(L.gz_gapped_extender_set_queries(h, DELIM, qp.ptr))
    gz.check(L.gz_gapped_extender_set_database(h, dp.ptr))
    gz.check(L.gz_gapped_extender_set_score_parameters(h, dev["d_matrix"].ptr, dev["nl"], 11, 1,
                                                       52))
    s = gz.Stream()

    def run(reverse, order):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        d_q, d_d = gz.DeviceBuffer.from_array(a.astype(np.uint32)), \
            gz.DeviceBuffer.from_array(b.astype(np.uint32))
        d_s, d_c = gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n * 4)
        if order is None:
            gz.check(L.gz_gapped_extender_extend_one_side_score_only_cells_async(
                h, n, int(reverse), d_q.ptr, d_d.ptr, d_s.ptr, d_c.ptr, s.handle))
        else:
            d_o = gz.DeviceBuffer.from_array(order.astype(np.uint32))
            gz.check(L.gz_gapped_extender_extend_one_side_score_only_ordered_async(
                h, n, int(reverse), d_o.ptr, d_q.ptr, d_d.ptr, d_s.ptr, d_c.ptr, s.handle))
        s.sync()
        return [x.to_array(t, n) for x, t in ((d_s, np.int32), (d_q, np.uint32),
                                               (d_d, np.uint32), (d_c, np.uint32))]

    def failures():
        f = (ctypes.c_uint32 * 1)()
        gz.check(L.gz_gapped_extender_failures(h, f))
        return f[0]

    for reverse in (False, True):
        want = run(reverse, None)
        for persistent in (0, 1):
            # the static grid and the persistent work-queue launch
            gz.check(L.gz_gapped_extender_set_persistent(h, persistent))
            for order in (rng.permutation(n), np.arange(n)[::-1]):
                got = run(reverse, order)
                for g, w in zip(got, want):
                    assert np.array_equal(g, w)
            got = run(reverse, None)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
            gz.check(L.gz_gapped_extender_set_persistent(h, 0))
        assert failures() == 0
    for persistent in (0, 1):
        gz.check(L.gz_gapped_extender_set_persistent(h, persistent))
        bad = rng.permutation(n)
        bad[7] = n + 5  # out of range: that slot is skipped and counted
        got = run(False, bad)
        assert failures() == 1
    L.gz_gapped_extender_destroy(h)


def test_gapped_kernel_wide_band(dev):
    """cutoff = INT_MAX disables X-drop: bands span whole sequences, so the
    LDS ring overflows and the wide-ring pass must produce the answers."""
    rng = np.random.default_rng(21)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 200, min_len=60, max_len=300)
    say("wide gapped seeds", qpos.size)
    for reverse in (False, True):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        hs, hq, hd = _gapped_gpu(dev, q, d, a, b, reverse, INT_MAX)
        want = _gapped_oracle(q, d, a, b, reverse, INT_MAX, dev["matrix"], dev["nl"])
        got = np.stack([hs, hq.astype(np.int64), hd.astype(np.int64)], 1)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("cutoff,gap_open,gap_ext", [(26, 11, 1), (52, 11, 1), (300, 5, 2),
                                                     (1023, 11, 1), (40, 0, 1), (35, 8, 0)])
def test_gapped_wave_per_band_vs_oracle(dev, monkeypatch, cutoff, gap_open, gap_ext):
    """One wave per band (the global ring's jobs): with GZ_GAPPED_WAVE=direct
    every band that outgrows a 16-column first ring runs on the wave kernel,
    64 columns per pass.  Long homologous pairs with indels give bands of
    many passes; cutoffs up to 1023 (the largest for which the wave form is
    exact), gap opening 0 and gap extension 0 included.  Scores and end
    positions are the oracle's, in both directions, and equal to the lane
    form's (GZ_GAPPED_WAVE=never)."""
    rng = np.random.default_rng(70 + cutoff + gap_open + gap_ext)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 400, min_len=100, max_len=2500,
                                                identity=(0.3, 0.97), indel=0.06)
    say("wave per band: cutoff", cutoff, "gaps", gap_open, gap_ext, "seeds", qpos.size)
    for reverse in (False, True):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        monkeypatch.setenv("GZ_GAPPED_WAVE", "direct")
        got = np.stack(_gapped_gpu(dev, q, d, a, b, reverse, cutoff, gap_open=gap_open,
                                   gap_ext=gap_ext, narrow=16), 1).astype(np.int64)
        monkeypatch.setenv("GZ_GAPPED_WAVE", "never")
        lane = np.stack(_gapped_gpu(dev, q, d, a, b, reverse, cutoff, gap_open=gap_open,
                                    gap_ext=gap_ext, narrow=16), 1).astype(np.int64)
        monkeypatch.delenv("GZ_GAPPED_WAVE")
        want = _gapped_oracle(q, d, a, b, reverse, cutoff, dev["matrix"], dev["nl"],
                              gap_open=gap_open, gap_ext=gap_ext)
        bad = np.nonzero((got != want).any(1))[0]
        assert bad.size == 0, (reverse, bad[:5], got[bad[:5]], want[bad[:5]])
        assert np.array_equal(lane, want)


def test_gapped_kernel_scores_past_16_bits(dev):
    """Identical 7,000-residue sequences seeded in the middle: each side
    scores past 16384, where the 16-bit first ring hands the band on to the
    32-bit ring; the answers stay the oracle's."""
    rng = np.random.default_rng(44)
    L = 7000
    seq = rng.integers(0, 20, L).astype(np.uint8)
    q = np.concatenate([[DELIM], seq, [DELIM]]).astype(np.uint8)
    d = np.concatenate([[DELIM], seq, [DELIM], rng.integers(0, 20, 300).astype(np.uint8),
                        [DELIM]]).astype(np.uint8)
    qpos = np.array([3500, 3400, 6990, 20], np.uint32)
    dpos = np.array([3500, 3400, 6990, 20], np.uint32)
    for reverse in (False, True):
        a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
        hs, hq, hd = _gapped_gpu(dev, q, d, a, b, reverse, 52)
        want = _gapped_oracle(q, d, a, b, reverse, 52, dev["matrix"], dev["nl"])
        got = np.stack([hs, hq.astype(np.int64), hd.astype(np.int64)], 1)
        say("past 16 bits: reverse", reverse, "scores", list(hs))
        assert got[:2, 0].min() > 16384
        assert np.array_equal(got, want)


def test_gapped_kernel_reference_gpu_test_answers(dev):
    """test/gapped_extender_gpu_test.cu:40-253 on the DNA test alphabet."""
    D = {"A": 0, "C": 1, "G": 2, "T": 3}
    mat = np.array([[2 if i == j else -1 for j in range(4)] for i in range(4)], np.int32).ravel()

    def enc(s):
        return np.array([5] + [D[c] for c in s] + [5], np.uint8)

    s1 = enc("AGCAG")
    hs, hq, hd = _gapped_gpu(dev, enc("AGCAC"), s1, np.array([1], np.uint32),
                             np.array([1], np.uint32), False, 0, 0, 1, mat, 4, 5)
    assert (hs[0], hq[0], hd[0]) == (8, 4, 4)
    hs, hq, hd = _gapped_gpu(dev, enc("AGTCAC"), s1, np.array([1], np.uint32),
                             np.array([1], np.uint32), False, INT_MAX, 0, 1, mat, 4, 5)
    assert (hs[0], hq[0], hd[0]) == (7, 5, 4)


def test_gapped_kernel_vs_golden_fixture(dev):
    """GappedKernel against the reference's own GappedExtender outputs
    (tests/golden/gapped_extender.npz)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "gapped_extender.npz"))
    q, d = g["query"], g["database"]
    for reverse in (0, 1):
        for cutoff in np.unique(g["cutoff"]):
            sel = np.nonzero((g["reverse"] == reverse) & (g["cutoff"] == cutoff))[0]
            say("golden gapped", reverse, int(cutoff), sel.size)
            hs, hq, hd = _gapped_gpu(dev, q, d, g["qstart"][sel].astype(np.uint32),
                                     g["dstart"][sel].astype(np.uint32), reverse, int(cutoff))
            assert np.array_equal(hs, g["score"][sel])
            assert np.array_equal(hq.astype(np.int64), g["qstart"][sel] + g["p0"][sel])
            assert np.array_equal(hd.astype(np.int64), g["dstart"][sel] + g["p1"][sel])


def _tb_jobs(rng, n, **kw):
    q, qo, d, do, qpos, dpos = related_pair_set(rng, n, **kw)
    jobs = np.zeros((n, 4), np.uint32)
    for i in range(n):
        lead = int(qo[i]) - 1
        jobs[i] = (lead, int(qo[i + 1] - qo[i]) + 1, int(qpos[i]) - lead, int(dpos[i]))
    return q, d, jobs, int(np.diff(do.astype(np.int64)).max())


@pytest.mark.parametrize("cutoff", [52, 26, 8])
def test_result_build_kernel_vs_host(dev, cutoff):
    """gz_build_results (one lane per result, the reference's gap_states_
    layout) equals the host result build on homologous pairs with indels:
    the same stale flag always, every field when no stale byte was read."""
    rng = np.random.default_rng(40 + cutoff)
    q, d, jobs, dmax = _tb_jobs(rng, 1200, min_len=5, max_len=400, identity=(0.25, 0.95),
                                indel=0.05)
    say("result build jobs", len(jobs), "cutoff", cutoff)
    _, pq = gz.pack_on_device(q)
    _, pd = gz.pack_on_device(d)
    out = np.zeros((len(jobs), 10), np.uint32)
    qlens = jobs[:, 1]
    gz.check(gz.lib().gz_build_results(pq.ptr, pd.ptr, gz.ptr(jobs), len(jobs), int(qlens.max()),
                                       dmax, dev["d_matrix"].ptr, dev["nl"], 11, 1, cutoff,
                                       gz.ptr(out), None))
    host = np.zeros(10, np.uint32)
    overflow = stale = 0
    for k, (lead, qlen, hq, hd) in enumerate(jobs):
        qs = np.ascontiguousarray(q[lead:lead + qlen])
        gz.check(gz.lib().gz_build_result_host(gz.ptr(qs), int(qlen), gz.ptr(d), int(hq), int(hd),
                                               cutoff, gz.ptr(host)))
        if out[k, 9] & 2:
            overflow += 1
            continue
        # a stale read makes the result history-dependent: both builds must flag
        # it (the caller then replays the chunk on the host); otherwise every
        # field is equal
        assert out[k, 9] == host[9], (k, out[k], host)
        if host[9]:
            stale += 1
            continue
        assert np.array_equal(out[k], host), (k, out[k], host)
    say("  overflow", overflow, "stale", stale)
    assert overflow < len(jobs) // 20


def test_result_build_kernel_large_cutoff(dev):
    """A cutoff above 8192 keeps the traceback's LDS columns in 32 bits: short
    query sides (at most 62 residues, the LDS path) with X-drop nearly off
    still equal the host result build."""
    rng = np.random.default_rng(47)
    q, d, jobs, dmax = _tb_jobs(rng, 400, min_len=5, max_len=60, identity=(0.25, 0.95),
                                indel=0.05)
    cutoff = 10000
    _, pq = gz.pack_on_device(q)
    _, pd = gz.pack_on_device(d)
    out = np.zeros((len(jobs), 10), np.uint32)
    gz.check(gz.lib().gz_build_results(pq.ptr, pd.ptr, gz.ptr(jobs), len(jobs),
                                       int(jobs[:, 1].max()), dmax, dev["d_matrix"].ptr,
                                       dev["nl"], 11, 1, cutoff, gz.ptr(out), None))
    host = np.zeros(10, np.uint32)
    for k, (lead, qlen, hq, hd) in enumerate(jobs):
        qs = np.ascontiguousarray(q[lead:lead + qlen])
        gz.check(gz.lib().gz_build_result_host(gz.ptr(qs), int(qlen), gz.ptr(d), int(hq), int(hd),
                                               cutoff, gz.ptr(host)))
        assert not out[k, 9] & 2, (k, out[k])
        assert out[k, 9] == host[9], (k, out[k], host)
        if not host[9]:
            assert np.array_equal(out[k], host), (k, out[k], host)
