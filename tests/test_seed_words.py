"""The seed stage's word-at-a-time tests (ghostz-gpu_amd/csrc/ungapped.h:
ungapped X-drop with trigger and the reduced-alphabet distance, twelve packed
codes per step) run on the host against the oracle's per-residue
restatements (oracle/gz_oracle.cpp, SeedSearcher::UngappedExtend and
DistanceCalculator::CalculateDistance), bit for bit.  CPU only."""
import os
import sys

import numpy as np
import pytest

if __name__ == "__main__":  # run by test_host_sanitizer.py as a child process
    _ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for _sub in ("ghostz-gpu_amd", "oracle"):
        sys.path.insert(0, os.path.join(_ROOT, _sub))

import gz_native as gz  # noqa: E402
import oracle_py  # noqa: E402

DELIM = 25


def pack(codes, tail=4):
    """The device's padded packed layout: 48 delimiter codes in front, 12
    five-bit codes per 64-bit word, four zero words behind (`tail`)."""
    padded = np.concatenate([np.full(48, DELIM, np.uint64), codes.astype(np.uint64)])
    nwords = (len(padded) + 11) // 12 + tail
    full = np.zeros(nwords * 12, np.uint64)
    full[:len(padded)] = padded
    shifts = (np.arange(12, dtype=np.uint64) * np.uint64(5))
    return np.bitwise_or.reduce(full.reshape(-1, 12) << shifts, axis=1).astype(np.uint64)


def sequences(rng, n, mean_len):
    """Delimiter-separated random residues (BLOSUM62 letters incl. BZJ*X),
    starting and ending with a delimiter, some runs of homologous text."""
    parts = [np.array([DELIM], np.uint8)]
    while sum(len(p) for p in parts) < n:
        L = int(rng.integers(1, 2 * mean_len))
        s = rng.integers(0, 25, L).astype(np.uint8)
        parts += [s, np.array([DELIM], np.uint8)]
    return np.concatenate(parts)


def related(rng, q, identity):
    d = q.copy()
    flip = rng.random(len(d)) > identity
    d[flip] = rng.integers(0, 25, int(flip.sum()))
    d[q == DELIM] = DELIM
    return d


def _ungapped_case(seed, identity, cut_hi):
    rng = np.random.default_rng(seed)
    q = sequences(rng, 4000, 40)
    d = related(rng, q, identity) if identity > 0 else sequences(rng, 4000, 60)
    n = 6000
    # seeds sit after the leading delimiter (position 0 is never a seed)
    qpos = rng.integers(1, len(q), n).astype(np.uint32)
    dpos = (qpos if identity > 0 else rng.integers(1, len(d), n)).astype(np.uint32)
    # positions right after / at delimiters, and every offset in a word
    dl = np.flatnonzero(q == DELIM).astype(np.uint32)
    dl = dl[dl > 0]
    k = min(200, len(dl))
    qpos[:k] = np.minimum(dl[:k] + 1, len(q) - 1)
    qpos[k:2 * k] = dl[:k]
    if identity > 0:
        dpos[:2 * k] = qpos[:2 * k]
        # the database side shifted by every offset against the query
        shift = np.arange(n - 2 * k) % 24 - 12
        dpos[2 * k:] = np.clip(qpos[2 * k:].astype(np.int64) + shift, 1, len(d) - 1)
    cut = rng.integers(1, cut_hi + 1, n).astype(np.int32)
    cut[:24] = np.arange(24) % 2 * (cut_hi - 1) + 1   # both ends of the range
    trig = rng.integers(0, 60, n).astype(np.int32)
    return q, d, qpos, dpos, cut, trig


def _oracle_ungapped(q, d, qpos, dpos, cut, trig, m, nl):
    O = oracle_py.oracle()
    return np.array([O.gzo_ungapped_with_trigger(gz.ptr(q), int(a), gz.ptr(d), int(b), DELIM,
                                                 gz.ptr(m), nl, int(c), int(t))
                     for a, b, c, t in zip(qpos, dpos, cut, trig)], np.int32)


@pytest.mark.parametrize("table_bytes,cut_hi", [(1, 40), (1, 128), (2, 40), (2, 3000)])
@pytest.mark.parametrize("seed,identity", [(1, 0.0), (2, 0.6), (3, 0.9), (4, 1.0)])
def test_ungapped_words_match_oracle(seed, identity, table_bytes, cut_hi):
    """Byte tables for cutoffs up to 128, 16-bit tables beyond (up to 16384)."""
    m, nl = gz.blosum62()
    q, d, qpos, dpos, cut, trig = _ungapped_case(seed, identity, cut_hi)
    qw, dw = pack(q), pack(d)
    n = len(qpos)
    got = np.zeros(n, np.int32)
    H = oracle_py.hostpipe()
    H.gzh_ungapped_words(gz.ptr(qw), gz.ptr(dw), n, gz.ptr(qpos), gz.ptr(dpos), gz.ptr(m), nl,
                         gz.ptr(cut), gz.ptr(trig), table_bytes, gz.ptr(got))
    want = _oracle_ungapped(q, d, qpos, dpos, cut, trig, m, nl)
    assert np.array_equal(got, want)
    # the cases reach well past one 12-residue word
    assert (want > 0).mean() > 0.3 or identity == 0.0
    if identity >= 0.9:
        assert want.max() > 100


def test_ungapped_words_asymmetric_matrix():
    """An asymmetric matrix tells the two tables (ba, ab) apart: a swapped
    row/column in either would change the scores."""
    m, nl = gz.blosum62()
    m = m.copy()
    rng = np.random.default_rng(9)
    mm = m.reshape(nl, nl)
    mm[:20, :20] += rng.integers(-2, 3, (20, 20)).astype(mm.dtype)
    m = mm.reshape(-1).copy()
    assert not np.array_equal(mm, mm.T)
    q, d, qpos, dpos, cut, trig = _ungapped_case(5, 0.8, 40)
    qw, dw = pack(q), pack(d)
    n = len(qpos)
    want = _oracle_ungapped(q, d, qpos, dpos, cut, trig, m, nl)
    for tb in (1, 2):
        got = np.zeros(n, np.int32)
        oracle_py.hostpipe().gzh_ungapped_words(gz.ptr(qw), gz.ptr(dw), n, gz.ptr(qpos),
                                                gz.ptr(dpos), gz.ptr(m), nl, gz.ptr(cut),
                                                gz.ptr(trig), tb, gz.ptr(got))
        assert np.array_equal(got, want)
    # the transposed matrix scores differently on these cases
    mt = mm.T.reshape(-1).copy()
    assert not np.array_equal(_oracle_ungapped(q, d, qpos, dpos, cut, trig, mt, nl), want)


@pytest.mark.parametrize("length", [4, 10, 12, 13, 25])
def test_distance_words_match_oracle(length):
    rng = np.random.default_rng(100 + length)
    rmap, _ = gz.reduced_map()
    q = sequences(rng, 3000, 50)
    d = related(rng, q, 0.7)
    qw, dw = pack(q), pack(d)
    n = 4000
    half = length // 2
    qc = rng.integers(half, len(q) - length, n).astype(np.uint32)
    dc = np.where(rng.random(n) < 0.5, qc, rng.integers(half, len(d) - length, n)).astype(np.uint32)
    got = np.zeros(n, np.int32)
    oracle_py.hostpipe().gzh_distance_words(gz.ptr(qw), gz.ptr(dw), n, gz.ptr(qc), gz.ptr(dc),
                                            length, gz.ptr(rmap), gz.ptr(got))
    O = oracle_py.oracle()
    want = np.array([O.gzo_distance(gz.ptr(q), int(a), gz.ptr(d), int(b), length, gz.ptr(rmap))
                     for a, b in zip(qc, dc)], np.int32)
    assert np.array_equal(got, want)
    assert len(set(want.tolist())) > 3


def run_bounds_cases(lib_path):
    """Every seed of sequences laid out as on the device but with only two
    tail words (the walk reads at most two words past the last delimiter's
    word), both table widths: run under AddressSanitizer, a read past the
    arrays aborts."""
    m, nl = gz.blosum62()
    H = oracle_py.hostpipe(lib_path)
    for seed, identity in [(11, 1.0), (12, 0.7), (13, 0.0)]:
        rng = np.random.default_rng(seed)
        q = sequences(rng, 300, 20)
        d = related(rng, q, identity) if identity > 0 else sequences(rng, 300, 30)
        qw, dw = pack(q, tail=2).copy(), pack(d, tail=2).copy()
        qpos = np.arange(1, len(q), dtype=np.uint32)
        dpos = np.clip(qpos, 1, len(d) - 1).astype(np.uint32)
        n = len(qpos)
        cut = np.full(n, 100, np.int32)
        trig = np.full(n, 1000, np.int32)
        for tb in (1, 2):
            got = np.zeros(n, np.int32)
            H.gzh_ungapped_words(gz.ptr(qw), gz.ptr(dw), n, gz.ptr(qpos), gz.ptr(dpos),
                                 gz.ptr(m), nl, gz.ptr(cut), gz.ptr(trig), tb, gz.ptr(got))
    print("bounds cases done")


if __name__ == "__main__":
    run_bounds_cases(sys.argv[1])
