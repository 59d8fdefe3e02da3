"""End-to-end search on the GPU against the CPU oracle.

The product (gz_aligner_align: seed search, ungapped and gapped extension on
the device, chain filter / AddResults / BuildResults / WriteOutput on the
host) must write byte-identical tabular output to the oracle's gzo_align on
the same prepared queries and database (whose preparation is checked against
the reference's own code in test_host_prep.py).
"""
import os
import sys

import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def matrix():
    gz.check(gz.lib().gz_set_device(0))
    return gz.blosum62()


def query_chunks(view, size):
    """Aligner query chunking: add queries while the framed total stays below size."""
    lens = np.diff(view["offsets"].astype(np.int64))
    out, i, n = [], 0, len(lens)
    while i < n:
        s, j = 0, i
        while j < n and s + lens[j] < size:
            s += lens[j]
            j += 1
        j = max(j, i + 1)
        out.append((i, j))
        i = j
    return out


def sub_view(v, a, b):
    o = v["offsets"]
    return {
        "number_queries": b - a,
        "codes": v["codes"][o[a]:o[b]],
        "offsets": (o[a:b + 1] - o[a]).astype(np.uint64),
        "names": v["names"][a:b],
        "lambda": v["lambda"][a:b],
        "K": v["K"][a:b],
        "translated": v["translated"],
        "dna_lengths": v["dna_lengths"][a:b] if v["translated"] else None,
    }


def oracle_text(qv, db, m, nl, chunk_size=1 << 62, **kw):
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    parts, stats = [], []
    for a, b in query_chunks(qv, chunk_size):
        t, st = oracle_py.align(sub_view(qv, a, b), chunks, info, m, nl, **kw)
        parts.append(t)
        stats.append(st)
    total = {k: sum(s[k] for s in stats) for k in stats[0]}
    return b"".join(parts), total


def run_case(name, queries, db, m, nl, schedules=(gz.SCHEDULE_PLAIN,), chunk_size=None, **params):
    say("case", name)
    qv = queries.view()
    kw = {}
    if chunk_size:
        params["queries_chunk_size"] = chunk_size
        kw["chunk_size"] = chunk_size
    for k in ("max_results", "max_results_per_subject"):
        if k in params:
            kw[k] = params[k]
    want, ost = oracle_text(qv, db, m, nl, **kw)
    say("  oracle lines", want.count(b"\n"), ost)
    assert want.count(b"\n") > 0 or name.startswith("edge")
    # the batched form with its tuned gapped layout, and with every task's
    # jobs in one launch per side (a layout the tuner reaches only after its
    # trials)
    forms = [(sch, 0, None) for sch in schedules] + [(schedules[-1], 1, None),
                                                      (schedules[-1], 1, "whole")]
    for sch, batched, layout in forms:
        old = os.environ.get("GZ_GAPPED_LAYOUT")
        if layout:
            os.environ["GZ_GAPPED_LAYOUT"] = layout
        try:
            al = gz.Aligner(db, schedule=sch, batched=batched, pipeline_blocks=6, **params)
        finally:
            if layout:
                if old is None:
                    del os.environ["GZ_GAPPED_LAYOUT"]
                else:
                    os.environ["GZ_GAPPED_LAYOUT"] = old
        if sch == gz.SCHEDULE_LEARNED:
            al.set_cost_model(gz.Mlp.init_weights(7))
        got, st = al.align(queries)
        say("  schedule", sch, "batched", batched, "lines", got.count(b"\n"), "hits",
            st["seed_hits"],
            "chained", st["chained_hits"], "gapped", st["gapped_extensions"])
        if got != want:
            gl, wl = got.split(b"\n"), want.split(b"\n")
            for i, (g, w) in enumerate(zip(gl, wl)):
                if g != w:
                    raise AssertionError(f"{name}: line {i} differs\n got {g}\nwant {w}")
            raise AssertionError(f"{name}: {len(gl)} vs {len(wl)} lines")
        assert st["seed_hits"] == ost["hits"]
        assert st["chained_hits"] == ost["hits_after_chain_filter"]
        assert st["gapped_extensions"] == ost["gapped_extensions"]
    return st, ost


def test_protein_search_matches_oracle(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261202, 3000, 80)
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do)
    run_case("protein", q, db, m, nl,
             schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LENGTH_SORTED, gz.SCHEDULE_LEARNED))


def test_translated_search_matches_oracle(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.translated_workload(20261203, 2000, 800)
    q = gz.Queries(qn, qr, qo, translated=True)
    db = gz.Database(dn, dr, do)
    run_case("translated", q, db, m, nl, schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LEARNED))


def test_unclustered_database_matches_oracle(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261204, 1500, 40)
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do, clustering=False)
    run_case("unclustered", q, db, m, nl)


def test_many_chunks_match_oracle(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261205, 2500, 60)
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do, chunk_size=200_000)
    assert db.info().number_chunks >= 3
    run_case("chunks", q, db, m, nl, chunk_size=4000)


def test_uneven_queries_match_oracle(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.uneven_protein_workload(20261206, 1200, 40,
                                                                   planted_fraction=0.5)
    lens = np.diff(qo.astype(np.int64))
    say("uneven query lengths: max", lens.max(), "median", np.median(lens))
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do)
    run_case("uneven", q, db, m, nl,
             schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LENGTH_SORTED, gz.SCHEDULE_LEARNED))


def test_result_limits_and_no_filter(matrix):
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261207, 1500, 30, homologs=8)
    q = gz.Queries(qn, qr, qo, filter=False)
    db = gz.Database(dn, dr, do)
    run_case("limits", q, db, m, nl, max_results=3, max_results_per_subject=2)


def long_protein_case():
    """Queries of 2,000-5,000 residues that are near copies of database
    subjects (scores far past 16 bits), one distant homolog, one unrelated."""
    rng = np.random.default_rng(20261209)
    db = synthetic.protein_database(rng, 400)
    long_subjects = [synthetic.random_codes(rng, int(n)) for n in rng.integers(2000, 5000, 8)]
    db = db + long_subjects
    queries = [synthetic.substitute(rng, s, 0.99) for s in long_subjects[:6]]
    queries += [synthetic.homolog(rng, long_subjects[6], 0.9), synthetic.random_codes(rng, 3000)]
    return synthetic._pack(queries, "query_"), synthetic._pack(db, "subject_")


def test_long_near_identical_proteins_match_oracle(matrix):
    """Ungapped and gapped scores run past 16 bits: the gapped bands leave
    the 16-bit rings for the global ring, and the result build walks
    alignments thousands of residues long.  Every form gives the oracle's
    text."""
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = long_protein_case()
    run_case("long", gz.Queries(qn, qr, qo), gz.Database(dn, dr, do), m, nl,
             schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LEARNED))


def test_edge_queries_match_oracle(matrix):
    m, nl = matrix
    rng = np.random.default_rng(20261208)
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261208, 800, 4)
    db_seqs = [dr[int(do[i]):int(do[i + 1])] for i in range(len(dn))]
    texts = [b"M", b"MK", b"ACDEFGH", b"X" * 40, b"ACDEFGHIKLMNPQ", db_seqs[3], db_seqs[5][:60],
             b"W" * 30, qr[int(qo[0]):int(qo[1])], b"BZJ*ACDEFGHIKLMNBZ" * 3]
    texts += [bytes(synthetic.AMINO[c] for c in synthetic.random_codes(rng, int(n)))
              for n in (8, 9, 10, 11, 12, 13)]
    names = [f"edge_{i}" for i in range(len(texts))]
    offs = np.zeros(len(texts) + 1, np.uint64)
    np.cumsum([len(t) for t in texts], out=offs[1:])
    q = gz.Queries(names, b"".join(texts), offs)
    db = gz.Database(dn, dr, do)
    run_case("edge", q, db, m, nl)


def test_protein_search_at_full_small_config_matches_oracle(matrix):
    """The plain protein workload at its stated size: 500 queries of 100-500
    aa, each planted three times into a 20,000-protein database."""
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(20261210, 20000, 500)
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do)
    st, ost = run_case("protein500x20000", q, db, m, nl,
                       schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LEARNED))
    assert st["results"] > 1500


def test_seed_hit_buffer_growth_matches_oracle(matrix, monkeypatch):
    """A 64-seed first hit buffer: every seed launch outgrows it, grows it
    keeping the hits already found and redoes its range.  The output is the
    oracle's."""
    m, nl = matrix
    monkeypatch.setenv("GZ_SEED_HIT_CAPACITY", "64")
    (qn, qr, qo), (dn, dr, do) = synthetic.translated_workload(20261211, 1500, 600)
    q = gz.Queries(qn, qr, qo, translated=True)
    db = gz.Database(dn, dr, do, chunk_size=150_000)
    assert db.info().number_chunks >= 2
    st, _ = run_case("hit-buffer-growth", q, db, m, nl,
                     schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LEARNED))
    assert st["seed_hits"] > 64 * 20


def test_uneven_queries_at_the_batched_shape_match_oracle(matrix):
    """The uneven-length protein workload at a reduced size: 300 queries of
    Pareto-tailed lengths (30 to 35,000 aa, median 250, 20% planted) against
    5,000 proteins.  Long queries carry long homologs whose bands outgrow both
    LDS rings; every form, including the batched persistent one, gives the
    oracle's text."""
    m, nl = matrix
    (qn, qr, qo), (dn, dr, do) = synthetic.uneven_protein_workload(20261212, 5000, 300)
    lens = np.diff(qo.astype(np.int64))
    say("uneven 300 x 5000: max query", lens.max(), "median", np.median(lens))
    assert lens.max() == 35000
    q = gz.Queries(qn, qr, qo)
    db = gz.Database(dn, dr, do)
    run_case("uneven300x5000", q, db, m, nl,
             schedules=(gz.SCHEDULE_PLAIN, gz.SCHEDULE_LEARNED))
