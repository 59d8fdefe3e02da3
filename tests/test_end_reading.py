"""Parity hazard 1: the reference CPU path's presearch `end` is read from an
uninitialised PresearchedResult and accumulated with += over the hits of one
Extend call (src/aligner_presearch_thread.cpp:322, 363-364); the GPU path
sets it to the forward extension's best cell
(src/aligner_gpu_presearch_thread.cpp:969-973).  The oracle follows the GPU
path.  `end` only feeds AddResults (order at score ties, the -b cut and the
containment check), so the other reading can change which results a query
keeps and their order.  This counts the queries whose output lines change
when the oracle reads `end` as zero-initialised and accumulated.
"""
import ctypes

import numpy as np

import gz_native as gz
import oracle_py
import synthetic


def count_changed(qd, dd, translated=False):
    m, nl = gz.blosum62()
    q = gz.Queries(*qd, translated=translated)
    db = gz.Database(*dd)
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    L = oracle_py.oracle()
    L.gzo_set_end_reading.argtypes = [ctypes.c_int]
    L.gzo_set_end_reading.restype = None
    outs = []
    try:
        for reading in (0, 1):
            L.gzo_set_end_reading(reading)
            outs.append(oracle_py.align(q.view(), chunks, info, m, nl)[0])
    finally:
        L.gzo_set_end_reading(0)

    def by_query(text):
        d = {}
        for ln in text.splitlines():
            d.setdefault(ln.split(b"\t")[0], []).append(ln)
        return d
    a, b = by_query(outs[0]), by_query(outs[1])
    names = set(a) | set(b)
    return sum(a.get(n) != b.get(n) for n in names), len(q.view()["names"]), outs


def test_end_reading_changes_are_counted():
    qd, dd = synthetic.protein_workload(20261202, 3000, 120)
    changed, n, outs = count_changed(qd, dd)
    print(f"end reading: {changed} of {n} queries change")
    # the hazard is real: with many hits per query the accumulated ends
    # reorder ties and change which results are kept
    assert 0 < changed <= n
    assert outs[0] != outs[1]
