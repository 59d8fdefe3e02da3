"""The seed stage's word-at-a-time tests (ghostz-gpu_amd/csrc/ungapped.h:
ungapped X-drop with trigger and the reduced-alphabet distance, twelve packed
codes per step) run on the host against the oracle's per-residue
restatements (oracle/gz_oracle.cpp, SeedSearcher::UngappedExtend and
DistanceCalculator::CalculateDistance), bit for bit.  CPU only."""
import numpy as np
import pytest

import gz_native as gz
import oracle_py

DELIM = 25


def pack(codes):
    """The device's padded packed layout: 48 delimiter codes in front, 12
    five-bit codes per 64-bit word, four zero words behind."""
    padded = np.concatenate([np.full(48, DELIM, np.uint64), codes.astype(np.uint64)])
    nwords = (len(padded) + 11) // 12 + 4
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


@pytest.mark.parametrize("seed,identity", [(1, 0.0), (2, 0.6), (3, 0.9), (4, 1.0)])
def test_ungapped_words_match_oracle(seed, identity):
    rng = np.random.default_rng(seed)
    m, nl = gz.blosum62()
    q = sequences(rng, 4000, 40)
    d = related(rng, q, identity) if identity > 0 else sequences(rng, 4000, 60)
    qw, dw = pack(q), pack(d)
    n = 6000
    # seeds sit after the leading delimiter (position 0 is never a seed)
    qpos = rng.integers(1, len(q), n).astype(np.uint32)
    dpos = (qpos if identity > 0 else rng.integers(1, len(d), n)).astype(np.uint32)
    # positions right after / at delimiters
    dl = np.flatnonzero(q == DELIM).astype(np.uint32)
    dl = dl[dl > 0]
    k = min(200, len(dl))
    qpos[:k] = np.minimum(dl[:k] + 1, len(q) - 1)
    qpos[k:2 * k] = dl[:k]
    if identity > 0:
        dpos[:2 * k] = qpos[:2 * k]
    cut = rng.integers(1, 40, n).astype(np.int32)
    trig = rng.integers(0, 60, n).astype(np.int32)
    got = np.zeros(n, np.int32)
    H = oracle_py.hostpipe()
    H.gzh_ungapped_words(gz.ptr(qw), gz.ptr(dw), n, gz.ptr(qpos), gz.ptr(dpos), gz.ptr(m), nl,
                         gz.ptr(cut), gz.ptr(trig), gz.ptr(got))
    O = oracle_py.oracle()
    want = np.array([O.gzo_ungapped_with_trigger(gz.ptr(q), int(a), gz.ptr(d), int(b), DELIM,
                                                 gz.ptr(m), nl, int(c), int(t))
                     for a, b, c, t in zip(qpos, dpos, cut, trig)], np.int32)
    assert np.array_equal(got, want)
    # the cases reach well past one 12-residue word
    assert (want > 0).mean() > 0.3 or identity == 0.0
    if identity >= 0.9:
        assert want.max() > 100


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
