"""Query hash lists (SeedSearcherQueryParameters::Build with one thread,
src/seed_searcher_query_parameters.cpp:17-219, visited as
src/aligner_presearch_thread.cpp:140-152 does): the product's host builder
(query_lists.cpp) on the CPU and its device builder (query_lists_device.hip)
on the GPU, each array equal to the oracle's (oracle/gz_oracle.cpp
BuildHashLists)."""
import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic


def oracle_lists(q, db):
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    return oracle_py.query_lists(q.view(), chunks, info)


def concat_offsets(q):
    offs = np.asarray(q.view()["offsets"], np.int64)
    lens = offs[1:] - offs[:-1]
    coff = np.zeros(len(lens), np.int64)
    np.cumsum(lens[:-1] - 1, out=coff[1:])
    return coff


def assert_same(got, want, q):
    assert np.array_equal(got["hash"], want["hash"])
    assert np.array_equal(got["entry_offsets"], want["entry_offsets"])
    assert np.array_equal(got["entry_queries"], want["entry_queries"])
    local = got["entry_positions"].astype(np.int64) - concat_offsets(q)[got["entry_queries"]]
    assert np.array_equal(local, want["entry_positions"].astype(np.int64))
    assert np.array_equal(got["entry_checks"], want["entry_checks"])


def cases():
    # (name, queries, database, translated): several 64-query blocks and
    # 256-list groups, six-frame reads, proteins with short segments
    (qn, qr, qo), dd = synthetic.protein_workload(51, 300, 150)
    yield "protein", (qn, qr, qo), dd, False
    qd, dd2 = synthetic.translated_workload(52, 300, 700)
    yield "translated", qd, dd2, True
    texts = [b"M", b"MKV", b"X" * 40, b"ACDEFGHIKLMNPQRSTVWY" * 4, b"W" * 30,
             b"BZJ*ACDEFGHIKLMNBZ" * 3, qr[int(qo[0]):int(qo[1])], b"ACDEFGHIK*LMNPQRSTVW"]
    offs = np.zeros(len(texts) + 1, np.uint64)
    np.cumsum([len(t) for t in texts], out=offs[1:])
    yield "edge", ([f"e{i}" for i in range(len(texts))], b"".join(texts), offs), dd, False


@pytest.mark.parametrize("case", ["protein", "translated", "edge"])
def test_host_lists_match_oracle(case):
    for name, qd, dd, tr in cases():
        if name != case:
            continue
        q = gz.Queries(*qd, translated=tr)
        db = gz.Database(*dd)
        want = oracle_lists(q, db)
        got = gz.query_lists(q, db, on_device=False)
        assert len(want["hash"]) > (0 if name == "edge" else 256)
        assert_same(got, want, q)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["protein", "translated", "edge"])
def test_device_lists_match_oracle(case):
    gz.check(gz.lib().gz_set_device(0))
    for name, qd, dd, tr in cases():
        if name != case:
            continue
        q = gz.Queries(*qd, translated=tr)
        db = gz.Database(*dd)
        want = oracle_lists(q, db)
        print("lists", name, len(want["hash"]), "entries", len(want["entry_queries"]), flush=True)
        got = gz.query_lists(q, db, on_device=True)
        assert_same(got, want, q)


@pytest.mark.gpu
def test_device_lists_match_host_at_scale():
    """40,000 translated reads (~12M query residues, ~1.2M lists): the device
    build equals the host build array for array."""
    gz.check(gz.lib().gz_set_device(0))
    qd, dd = synthetic.translated_workload(53, 2000, 40000)
    q = gz.Queries(*qd, translated=True)
    db = gz.Database(*dd)
    host = gz.query_lists(q, db, on_device=False)
    dev = gz.query_lists(q, db, on_device=True)
    print("lists", len(host["hash"]), "entries", len(host["entry_queries"]),
          "host s", host["seconds"], "device s", dev["seconds"], flush=True)
    for k in ("hash", "entry_offsets", "entry_positions", "entry_queries", "entry_checks"):
        assert np.array_equal(host[k], dev[k]), k
