"""The on-disk database (<prefix>.inf, <prefix>_<i>.{inf,off,nam,seq,scp,sdp};
src/database.h:150-152, src/database_chunk.h:103-123) read by
gz_database_load and written by gz_database_save.

* Files in the reference's layout, with the .sdp / .scp streams written by
  the reference's own SeedSearcherDatabaseParameters::Save /
  SaveCommonParameters (oracle/_ref), load into exactly the index the
  product builds from the same sequences; the product's writer reproduces
  every file but .sdp byte for byte (.sdp holds unordered hash maps) and
  its .sdp loads back to the same index.
* The committed fixture tests/golden/db_test_protein (the reference test's
  test_protein.fa, chunked at 60 residues as in test/database_test.cpp:31)
  loads with the sizes test/database_test.cpp:97-118 expects.
* Missing, truncated and inconsistent files are errors naming the file.
CPU only."""
import os
import shutil

import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURE = os.path.join(GOLDEN, "db_test_protein")
ARRAYS = ("sequence", "offsets", "hashes", "rep_offsets", "rep_positions", "rep_member_offsets",
          "member_positions", "value_offsets", "value_positions")
INFO = ("number_chunks", "total_length", "total_sequences", "max_length", "subsequence_length",
        "hash_max_code", "hash_score_threshold", "hash_max_length", "hash_shift",
        "max_mismatches")

needs_ref = pytest.mark.skipif(not oracle_py.ref_available(),
                               reason="oracle/_ref is not built (the reference tree is absent)")


def same_database(a, b):
    ia, ib = a.info(), b.info()
    for k in INFO:
        assert getattr(ia, k) == getattr(ib, k), k
    assert bytes(ia.reduced_code_map) == bytes(ib.reduced_code_map)
    assert bytes(ia.hash_code_map) == bytes(ib.hash_code_map)
    for c in range(ia.number_chunks):
        ca, cb = a.chunk(c), b.chunk(c)
        for k in ARRAYS:
            assert np.array_equal(ca[k], cb[k]), (c, k)
        assert ca["names"] == cb["names"]
        assert ca["clustering"] == cb["clustering"]


def chunks_of(db):
    out = []
    for c in range(db.info().number_chunks):
        ch = db.chunk(c)
        out.append((ch["sequence"], ch["offsets"], ch["names"]))
    return out


def reference_files(tmp_path, db, clustering, L=10, s=39):
    m, nl = gz.blosum62()
    i = db.info()
    prefix = str(tmp_path / "ref")
    oracle_py.write_reference_database(prefix, chunks_of(db), i.max_length, i.total_length,
                                       i.total_sequences, clustering, L, s, m, nl)
    return prefix


@needs_ref
@pytest.mark.parametrize("clustering", [True, False])
def test_reference_files_load_into_the_built_index(tmp_path, clustering):
    (_, _, _), (dn, dr, do) = synthetic.protein_workload(20261215, 120, 10, qlen=(100, 300),
                                                         homologs=12)
    built = gz.Database(dn, dr, do, clustering=clustering, chunk_size=12_000)
    assert built.info().number_chunks >= 3
    prefix = reference_files(tmp_path, built, clustering)
    loaded = gz.Database.load(prefix)
    same_database(loaded, built)
    if clustering:
        assert sum(loaded.chunk(c)["member_positions"].size
                   for c in range(loaded.info().number_chunks)) > 0


@needs_ref
def test_writer_reproduces_the_reference_files(tmp_path):
    (_, _, _), (dn, dr, do) = synthetic.protein_workload(20261216, 60, 10, qlen=(100, 300),
                                                         homologs=8)
    built = gz.Database(dn, dr, do, chunk_size=8_000)
    ref = reference_files(tmp_path, built, True)
    mine = str(tmp_path / "mine")
    built.save(mine)
    n = built.info().number_chunks
    for name in ["ref.inf"] + [f"ref_{c}.{e}" for c in range(n)
                               for e in ("inf", "off", "nam", "seq", "scp")]:
        a = (tmp_path / name).read_bytes()
        b = (tmp_path / name.replace("ref", "mine", 1)).read_bytes()
        assert a == b, name
    same_database(gz.Database.load(mine), built)


def test_save_load_round_trip(tmp_path):
    (_, _, _), (dn, dr, do) = synthetic.protein_workload(20261217, 80, 10, qlen=(100, 300),
                                                         homologs=8)
    for clustering in (True, False):
        built = gz.Database(dn, dr, do, clustering=clustering, chunk_size=10_000)
        prefix = str(tmp_path / f"db{int(clustering)}")
        built.save(prefix)
        same_database(gz.Database.load(prefix), built)


def test_fixture_matches_the_reference_database_test(tmp_path):
    db = gz.Database.load(os.path.join(FIXTURE, "test_protein"))
    i = db.info()
    assert i.number_chunks == 4                    # GetNumberChunks
    assert i.total_length == 198                   # GetDatabaseTotalLenght
    assert i.total_sequences == 8                  # GetNumberTotalSequences
    assert i.max_length == 25                      # GetMaxSequenceLength
    c0 = db.chunk(0)
    assert len(c0["names"]) == 2                   # GetNumberSequencesInChunk
    assert c0["sequence"].size == 53               # GetConcatenatedSequenceLength
    assert c0["names"] == [b"test0", b"test1"]
    built = gz.Database.from_fasta(os.path.join(FIXTURE, "test_protein.fa"), chunk_size=60)
    same_database(db, built)
    assert sum(db.chunk(c)["member_positions"].size for c in range(4)) > 0


def test_missing_truncated_and_inconsistent_files(tmp_path):
    src = os.path.join(FIXTURE, "test_protein")
    work = tmp_path / "w"

    def fresh():
        if work.exists():
            shutil.rmtree(work)
        shutil.copytree(FIXTURE, work)
        return str(work / "test_protein")

    def fails(prefix, what):
        with pytest.raises(gz.GzError) as e:
            gz.Database.load(prefix)
        assert what in str(e.value), str(e.value)

    fails(str(tmp_path / "absent"), "absent.inf: cannot open")
    p = fresh()
    os.remove(p + "_2.sdp")
    fails(p, "test_protein_2.sdp: cannot open")
    p = fresh()
    data = open(p + "_1.sdp", "rb").read()
    open(p + "_1.sdp", "wb").write(data[:len(data) // 2])
    fails(p, "test_protein_1.sdp")
    p = fresh()
    data = bytearray(open(p + "_0.seq", "rb").read())
    data[0] = 3  # the leading delimiter
    open(p + "_0.seq", "wb").write(bytes(data))
    fails(p, "chunk sequence and offsets do not match")
    p = fresh()
    open(p + "_3.nam", "w").write("only_one\n")
    fails(p, "fewer names than sequences")
    assert os.path.exists(src + ".inf")


@pytest.mark.gpu
def test_gpu_align_main_on_fixture_matches_fasta_database(tmp_path):
    """The search over the loaded fixture files is the search over the
    database built from its FASTA."""
    gz.check(gz.lib().gz_set_device(0))
    fa = os.path.join(FIXTURE, "test_protein.fa")
    p = gz.default_parameters()
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    gz.check(gz.lib().gz_align_main(fa.encode(), os.path.join(FIXTURE, "test_protein").encode(),
                                    a.encode(), 0, p))
    dbo = gz.DatabaseOptions(1, 10, 39, 60, 0)
    gz.check(gz.lib().gz_align_fasta(fa.encode(), fa.encode(), b.encode(), 0, p, dbo))
    ta, tb = open(a, "rb").read(), open(b, "rb").read()
    assert ta == tb
    # the host flow with the device stages emulated by the oracle
    from test_readme_example import read_fasta
    d = read_fasta(fa)
    want, _ = oracle_py.host_align(d, d, p, translated=False, db_chunk_size=60)
    assert ta == want
    assert ta.count(b"\n") == 11
