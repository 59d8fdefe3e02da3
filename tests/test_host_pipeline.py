"""The product's host flow on the CPU (no GPU): AlignQueries with its two
device stages emulated by the oracle's leaf functions (oracle/libhostpipe.so).

Checks query chunking, hash-list order, per-query seed order, chain filter,
gapped schedules, AddResults, BuildResults and WriteOutput against the
oracle's whole-search restatement, byte for byte.
"""
import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic
from test_pipeline_gpu import long_protein_case, oracle_text


def check(name, qd, dd, translated=False, filter=True, clustering=True, db_chunk_size=0,
          schedules=(0, 1), batched=(0, 1), force_stale=0, **params):
    m, nl = gz.blosum62()
    q = gz.Queries(*qd, translated=translated, filter=filter)
    db = gz.Database(*dd, clustering=clustering, chunk_size=db_chunk_size)
    kw = {}
    if "queries_chunk_size" in params:
        kw["chunk_size"] = params["queries_chunk_size"]
    ok = {k: params[k] for k in ("max_results", "max_results_per_subject") if k in params}
    want, ost = oracle_text(q.view(), db, m, nl, **kw, **ok)
    for sch in schedules:
        for bat in batched:
            p = gz.default_parameters(schedule=sch, record_traces=1, batched=bat,
                                      pipeline_blocks=5, **params)
            got, st = oracle_py.host_align(qd, dd, p, translated=translated, filter=filter,
                                           clustering=clustering, db_chunk_size=db_chunk_size,
                                           force_stale=force_stale)
            assert got == want, (name, sch, bat)
            assert st["seed_hits"] == ost["hits"]
            assert st["chained_hits"] == ost["hits_after_chain_filter"]
            assert st["gapped_extensions"] == ost["gapped_extensions"]
    return want


def test_protein():
    w = check("protein", *synthetic.protein_workload(31, 1500, 40))
    assert w.count(b"\n") > 40


def test_long_near_identical_proteins():
    # alignments thousands of residues long, scores past 16 bits
    qd, dd = long_protein_case()
    check("long", qd, dd)


def test_translated():
    w = check("translated", *synthetic.translated_workload(32, 1000, 400), translated=True)
    assert w.count(b"\n") > 100


def test_unclustered_and_no_filter():
    check("unclustered", *synthetic.protein_workload(33, 800, 20), clustering=False,
          filter=False)


def test_query_and_database_chunks():
    qd, dd = synthetic.protein_workload(34, 1500, 40)
    check("chunks", qd, dd, db_chunk_size=120_000, queries_chunk_size=3000)


def test_result_limits():
    check("limits", *synthetic.protein_workload(35, 800, 20, homologs=8), max_results=3,
          max_results_per_subject=2)


@pytest.mark.parametrize("threads", [1, 3])
def test_thread_count_does_not_change_output(threads):
    check("threads", *synthetic.protein_workload(36, 800, 30), threads=threads, schedules=(0,))


def test_edge_queries():
    (qn, qr, qo), dd = synthetic.protein_workload(37, 400, 3)
    dn, dr, do = dd
    texts = [b"M", b"MK", b"X" * 40, dr[int(do[3]):int(do[4])], b"W" * 30,
             b"BZJ*ACDEFGHIKLMNBZ" * 3, qr[int(qo[0]):int(qo[1])]]
    offs = np.zeros(len(texts) + 1, np.uint64)
    np.cumsum([len(t) for t in texts], out=offs[1:])
    check("edge", ([f"e{i}" for i in range(len(texts))], b"".join(texts), offs), dd)


def test_many_blocks_and_database_chunks():
    # several 64-query pipeline blocks, each seeded, extended, built and
    # written on its own, over three database chunks
    qd, dd = synthetic.protein_workload(38, 1500, 330)
    check("blocks", qd, dd, db_chunk_size=150_000, schedules=(0,))


def test_stale_traceback_rebuilds_in_reference_order():
    # a stale device traceback makes the host rebuild and rewrite everything
    qd, dd = synthetic.protein_workload(39, 1200, 200)
    check("stale", qd, dd, db_chunk_size=200_000, schedules=(0,), force_stale=11)


def test_small_many_blocks():
    # five 64-query blocks over small database chunks, with and without
    # stale tracebacks (small enough for the ThreadSanitizer run)
    qd, dd = synthetic.protein_workload(40, 150, 300)
    check("small blocks", qd, dd, db_chunk_size=20_000, schedules=(0,), batched=(1,))
    check("small blocks stale", qd, dd, db_chunk_size=20_000, schedules=(0,), batched=(1,),
          force_stale=5)


def test_small_blocks_two_extenders():
    # queries past 2048 residues: the batched form runs two extender threads
    # (even and odd tasks at once, published in task order)
    qd, dd = synthetic.protein_workload(41, 150, 70, qlen=(1800, 2300))
    assert np.diff(qd[2].astype(np.int64)).max() > 2048
    check("two extenders", qd, dd, db_chunk_size=20_000, schedules=(0,), batched=(1,))
