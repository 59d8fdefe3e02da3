"""The reference's one end-to-end example (README.md:88-97): data/db.fasta
formatted with the `db` defaults (-C T -L 10 -s 39) and data/queries.fasta
searched with `aln -q d -t p` (-b 10 -v 1 -F T).  The inputs and the eight
printed result lines are committed under tests/golden/readme_example/
(make_readme_example.py copied them).

This pins the oracle's whole-search composition -- query translation, hash
lists and their visiting order, chain filter, AddResults order and caps,
BuildResults and the output fields -- to output the reference itself
printed, and then the product to the same lines.
"""
import os

import numpy as np
import pytest

import gz_native as gz
import oracle_py

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "readme_example")
DB = os.path.join(HERE, "db.fasta")
QUERIES = os.path.join(HERE, "queries.fasta")


def expected_rows():
    with open(os.path.join(HERE, "expected.tsv")) as f:
        return [ln.split("\t") for ln in f.read().splitlines()]


def rows(text):
    """Fields of each output line as README.md shows them (tabs expanded, so
    white space separates fields).  Each line must have exactly 12
    tab-separated fields."""
    out = []
    for ln in text.decode().splitlines():
        assert len(ln.split("\t")) == 12, ln
        out.append(ln.split())
    return out


def check_names(text):
    # FastaSequenceReader::Read keeps a header's trailing blank in the name
    # (src/fasta_sequence_reader.cpp:41-47): ">query1 " names "query1 "
    names = [ln.split(b"\t")[0] for ln in text.splitlines()]
    assert b"query1 " in names and b"query1" not in names


def read_fasta(path):
    """(names, residues, offsets) as gz_queries_create takes them, read the
    way FastaSequenceReader::Read reads them: the name is the header after
    its leading '>' and blanks, a sequence line loses its outer white space."""
    names, seqs = [], []
    for ln in open(path):
        ln = ln.rstrip("\r\n")
        if ln.startswith(">"):
            names.append(ln.lstrip("> "))
            seqs.append([])
        elif ln:
            seqs[-1].append(ln.strip())
    seqs = ["".join(s) for s in seqs]
    offs = np.zeros(len(seqs) + 1, np.uint64)
    np.cumsum([len(s) for s in seqs], out=offs[1:])
    return names, "".join(seqs).encode(), offs


def test_expected_rows_shape():
    want = expected_rows()
    assert len(want) == 8 and all(len(r) == 12 for r in want)


def test_oracle_matches_reference_output():
    m, nl = gz.blosum62()
    q = gz.Queries.from_fasta(QUERIES, translated=True)
    db = gz.Database.from_fasta(DB)
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    got, st = oracle_py.align(q.view(), chunks, info, m, nl)
    assert rows(got) == expected_rows()
    check_names(got)
    assert st["hits"] > 0


def test_oracle_breaks_on_a_wrong_result_cap():
    """The example keeps two results for query0, query1, query3 and query4,
    so an oracle that capped -b one too low fails it.  (The example's
    alignments are gapless and exact, so it does not tell cutoffs or gap
    costs apart; the leaf tests pin those.)"""
    m, nl = gz.blosum62()
    q = gz.Queries.from_fasta(QUERIES, translated=True)
    db = gz.Database.from_fasta(DB)
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    want = expected_rows()
    got, _ = oracle_py.align(q.view(), chunks, info, m, nl, max_results=1)
    assert rows(got) != want


def test_host_flow_matches_reference_output():
    """The product's host flow on the CPU (device stages emulated) in both
    forms and every schedule."""
    qd, dd = read_fasta(QUERIES), read_fasta(DB)
    for sch in (gz.SCHEDULE_PLAIN, gz.SCHEDULE_LENGTH_SORTED):
        for bat in (0, 1):
            p = gz.default_parameters(schedule=sch, batched=bat)
            got, _ = oracle_py.host_align(qd, dd, p, translated=True)
            assert rows(got) == expected_rows(), (sch, bat)
            check_names(got)


@pytest.mark.gpu
def test_gpu_align_fasta_matches_reference_output(tmp_path):
    gz.check(gz.lib().gz_set_device(0))
    out = str(tmp_path / "exout")
    p = gz.default_parameters()
    dbo = gz.DatabaseOptions(1, 10, 39, 0, 0)
    gz.check(gz.lib().gz_align_fasta(QUERIES.encode(), DB.encode(), out.encode(), 1,
                                     p, dbo))
    text = open(out, "rb").read()
    assert rows(text) == expected_rows()
    check_names(text)


@pytest.mark.gpu
def test_gpu_forms_match_reference_output():
    gz.check(gz.lib().gz_set_device(0))
    q = gz.Queries.from_fasta(QUERIES, translated=True)
    db = gz.Database.from_fasta(DB)
    weights = gz.Mlp.init_weights(7)
    for sch in (gz.SCHEDULE_PLAIN, gz.SCHEDULE_LENGTH_SORTED, gz.SCHEDULE_LEARNED):
        for bat in (0, 1):
            al = gz.Aligner(db, schedule=sch, batched=bat)
            if sch == gz.SCHEDULE_LEARNED:
                al.set_cost_model(weights)
            got, _ = al.align(q)
            assert rows(got) == expected_rows(), (sch, bat)
            check_names(got)


@pytest.mark.gpu
def test_gpu_align_main_on_saved_database_matches_reference_output(tmp_path):
    """`aln -d <prefix>` on the database files (gz_database_save of the
    README database) gives the README's output."""
    gz.check(gz.lib().gz_set_device(0))
    prefix = str(tmp_path / "exdb")
    gz.Database.from_fasta(DB).save(prefix)
    out = str(tmp_path / "exout")
    gz.check(gz.lib().gz_align_main(QUERIES.encode(), prefix.encode(), out.encode(), 1,
                                    gz.default_parameters()))
    text = open(out, "rb").read()
    assert rows(text) == expected_rows()
    check_names(text)


def test_align_main_reports_missing_inputs(tmp_path):
    """AlignMain::Run's checks (src/align_main.cpp:46-62), before any device
    work."""
    out = str(tmp_path / "o").encode()
    p = gz.default_parameters()
    st = gz.lib().gz_align_main(str(tmp_path / "none.fa").encode(), b"x", out, 1, p)
    assert st == 4 and gz.lib().gz_last_error() == b"Query file does not exist."
    st = gz.lib().gz_align_main(QUERIES.encode(), str(tmp_path / "nodb").encode(), out, 1, p)
    assert st == 4 and gz.lib().gz_last_error() == b"Database file does not exist."
