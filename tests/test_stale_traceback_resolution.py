"""Results whose traceback reads op bytes that earlier extensions left.

The reference builds a database chunk's results through one extender, in
blocks of 32 results taken from the back of each block
(aligner_build_results_thread.h:97-150); its op arena
(src/gapped_extender.cpp:25-68) is never cleared, so a traceback that reads
outside its own extension's bytes gets what earlier extensions wrote.  The
product rebuilds only those results (ResolveStaleChunk): each such read is
answered by the latest earlier extension, in that order, that wrote the
byte.  The pins: the oracle's BuildResults of the chunk (gzo_build_chunk,
one extender in the reference's order) and the product's own sequential
replay of the chunk (ReplayChunkHost); all three must agree, on chunks
where many flagged results differ from a fresh extender's (so a resolver
that ignored the history, or took the wrong extension's byte, would fail).
CPU only (oracle/liboracle.so and the shim oracle/libhostpipe.so)."""
import numpy as np
import pytest

import gz_native as gz
import oracle_py
from helpers import DELIM, related_pair_set


def chunk_case(seed, pairs=60, hits_per_pair=3):
    rng = np.random.default_rng(seed)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, pairs, min_len=100, max_len=2000,
                                                identity=(0.2, 0.9), indel=0.08)
    starts = np.asarray(qo, np.uint64) - np.uint64(1)  # each query's leading delimiter
    hits = []
    for i in range(len(starts) - 1):
        for k in range(hits_per_pair):
            a, b = int(qpos[i]) + 7 * k, int(dpos[i]) + 7 * k
            if a < int(starts[i + 1]) and b < d.size - 1 and q[a] != DELIM and d[b] != DELIM:
                hits.append((i, a - int(starts[i]), b))
    return q, starts, d, np.array(hits, np.uint32)


@pytest.mark.parametrize("seed", [1, 4, 5, 9, 11])
def test_flagged_results_match_the_chunk_replay(seed):
    m, nl = gz.blosum62()
    m = np.asarray(m, np.int32)
    q, starts, d, hits = chunk_case(seed)
    (flagged, decided, differ, history), got = oracle_py.stale_resolution(
        q, starts, d, hits, m, nl, 11, 1, 30)
    assert flagged > 0 and history > 0, (flagged, history)
    assert decided == 1
    assert differ == 0
    want = oracle_py.build_chunk(q, starts, d, hits, m, nl, 11, 1, 30, DELIM)
    assert np.array_equal(got, want)


def test_totals_over_many_chunks():
    m, nl = gz.blosum62()
    tot_flagged = tot_history = 0
    for seed in range(12, 24):
        for cutoff in (30, 60):
            q, starts, d, hits = chunk_case(seed)
            (f, decided, differ, h), got = oracle_py.stale_resolution(
                q, starts, d, hits, np.asarray(m, np.int32), nl, 11, 1, cutoff)
            assert decided == 1 and differ == 0, (seed, cutoff, f, h)
            want = oracle_py.build_chunk(q, starts, d, hits, np.asarray(m, np.int32), nl, 11,
                                         1, cutoff, DELIM)
            assert np.array_equal(got, want), (seed, cutoff)
            tot_flagged += f
            tot_history += h
    assert tot_history >= 10, (tot_flagged, tot_history)
