"""Writes the golden fixtures of tests/golden/ from the reference's own leaf
code (oracle/_ref/libghostz_ref.so, compiled unmodified from the reference
sources by oracle/Makefile).  Inputs are seeded; every fixture stores its
inputs next to the reference outputs.

  python tests/golden/make_golden.py [fixture ...]   (needs oracle/_ref built;
                                                     all fixtures by default)
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("ghostz-gpu_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import gz_native  # noqa: E402
import oracle_py  # noqa: E402
from helpers import DELIM, related_pair_set, random_protein  # noqa: E402

R = None
REFERENCE = os.environ.get("GZ_REFERENCE", "/root/reference")


def ref_matrix():
    m = np.zeros(625, np.int32)
    nl = ctypes.c_uint32()
    R.gzr_read_matrix(gz_native.blosum62_text(), oracle_py.ptr(m), ctypes.byref(nl))
    return m, nl.value


def gapped(path):
    rng = np.random.default_rng(20261202)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 120, min_len=8, max_len=220,
                                                identity=(0.25, 0.95), indel=0.05)
    m, nl = ref_matrix()
    rows = []
    edits_all, edit_off = [], [0]
    for k in range(len(qpos)):
        for reverse in (0, 1):
            for cutoff in (26, 8, 52):
                a, b = (int(qpos[k]) - 1, int(dpos[k]) - 1) if reverse else (int(qpos[k]),
                                                                            int(dpos[k]))
                s, p0, p1, ed, _ = oracle_py.gapped_one_side("ref", q, a, q.size, d, b, DELIM,
                                                             reverse, m, nl, 11, 1, cutoff,
                                                             traceback=True)
                rows.append((a, b, q.size, reverse, cutoff, 11, 1, s, p0, p1))
                edits_all.extend(ed)
                edit_off.append(len(edits_all))
    r = np.array(rows, np.int64)
    np.savez_compressed(path, query=q, database=d, qstart=r[:, 0], dstart=r[:, 1],
                        seq0_len=r[:, 2], reverse=r[:, 3], cutoff=r[:, 4], gap_open=r[:, 5],
                        gap_ext=r[:, 6], score=r[:, 7], p0=r[:, 8], p1=r[:, 9],
                        edits=np.array(edits_all, np.uint8), edit_offsets=np.array(edit_off,
                                                                                    np.int64))


def ungapped_distance_hash(path):
    rng = np.random.default_rng(20261203)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 300, min_len=12, max_len=200)
    m, nl = ref_matrix()
    rmap = np.zeros(25, np.uint8)
    mx = ctypes.c_uint32()
    R.gzr_reduced_map(b"A KR EDNQ C G H ILVM FYW P ST", oracle_py.ptr(rmap), ctypes.byref(mx))
    ung = []
    for k in range(len(qpos)):
        for reverse in (0, 1):
            for cutoff in (15, 5):
                a, b = (int(qpos[k]) - 1, int(dpos[k]) - 1) if reverse else (int(qpos[k]),
                                                                            int(dpos[k]))
                b0, b1 = ctypes.c_int(), ctypes.c_int()
                s = R.gzr_ungapped_one_side_positions(
                    ctypes.c_void_p(q.ctypes.data + a), ctypes.c_void_p(d.ctypes.data + b),
                    DELIM, reverse, oracle_py.ptr(m), nl, cutoff, ctypes.byref(b0),
                    ctypes.byref(b1))
                ung.append((a, b, reverse, cutoff, s, b0.value, b1.value))
    dist = []
    for k in range(len(qpos)):
        la, lb = int(qo[k + 1] - qo[k] - 1), int(do[k + 1] - do[k] - 1)
        if la < 12 or lb < 12:
            continue
        for _ in range(4):
            a = int(rng.integers(qo[k] + 5, qo[k + 1] - 6))
            b = int(rng.integers(do[k] + 5, do[k + 1] - 6))
            dist.append((a, b, R.gzr_distance(oracle_py.ptr(q), a, oracle_py.ptr(d), b, 10,
                                              oracle_py.ptr(rmap))))
    scores = np.array([4, 5, 6, 9, 6, 8, 5, 11, 7, 5], np.int32)  # per reduced letter
    hashes = []
    seq = np.concatenate([[DELIM], random_protein(rng, 3000, ambiguous=0.02), [DELIM]]).astype(
        np.uint8)
    for i in range(1, seq.size - 1):
        h, hl = ctypes.c_uint32(), ctypes.c_uint32()
        ret = R.gzr_hash(ctypes.c_void_p(seq.ctypes.data + i), oracle_py.ptr(rmap), mx.value, 39,
                         8, oracle_py.ptr(scores), ctypes.byref(h), ctypes.byref(hl))
        hashes.append((i, ret, h.value, hl.value))
    np.savez_compressed(path, query=q, database=d, reduced_map=rmap,
                        max_regular=np.int64(mx.value), hash_scores=scores, hash_sequence=seq,
                        ungapped=np.array(ung, np.int64), distance=np.array(dist, np.int64),
                        hash=np.array(hashes, np.int64))


def statistics(path):
    rng = np.random.default_rng(20261204)
    m, nl = ref_matrix()
    seqs, offs, karlin = [], [0], []
    for trial in range(40):
        n = int(rng.integers(1, 400)) if trial % 5 else int(rng.integers(1, 6))
        body = random_protein(rng, n, ambiguous=0.02)
        if trial % 7 == 0:
            body[:] = body[0]
        s = np.concatenate([[DELIM], body, [DELIM]]).astype(np.uint8)
        lam, K, H = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ok = R.gzr_query_karlin(oracle_py.ptr(s), s.size, oracle_py.ptr(m), nl, ctypes.byref(lam),
                                ctypes.byref(K), ctypes.byref(H))
        karlin.append((ok, lam.value, K.value, H.value))
        seqs.append(s)
        offs.append(offs[-1] + s.size)
    space = []
    for ql in (1, 50, 150, 300, 35000):
        for dl, dn in ((1000, 3), (330_000_000, 1_000_000), (1_300_000_000, 4_000_000)):
            space.append((ql, dl, dn, R.gzr_search_space(ql, dl, dn)))
    conv = []
    for lam, K in ((0.267, 0.041), (0.3176, 0.134), (0.29, 0.11)):
        for x in (7.0, 15.0, 22.0, 25.0):
            conv.append((lam, K, x, R.gzr_normalized2nominal(x, lam, K),
                         R.gzr_normalized_cutoff2nominal(x, lam, K)))
    ev = []
    for s in (20, 52, 100, 400):
        for sp in (10**6, 10**11):
            ev.append((s, sp, R.gzr_nominal2evalue(s, sp, 0.267, 0.041),
                       R.gzr_nominal2normalized(s, 0.267, 0.041)))
    np.savez_compressed(path, sequences=np.concatenate(seqs), offsets=np.array(offs, np.int64),
                        karlin=np.array(karlin, np.float64), search_space=np.array(space, np.int64),
                        conversions=np.array(conv, np.float64), evalues=np.array(ev, np.float64))


def seg_translate(path):
    rng = np.random.default_rng(20261205)
    amino = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", np.uint8)
    prot, pmask, pstat, poff = [], [], [], [0]
    for n in list(range(0, 30)) + [60, 100, 300, 1000]:
        s = amino[rng.integers(0, 20, n)].copy()
        if n > 30 and rng.random() < 0.6:
            a = int(rng.integers(0, n - 20))
            s[a:a + 20] = s[a]
        b = ctypes.create_string_buffer(n + 1)
        st = R.gzr_seg(s.tobytes(), b)
        prot.append(s)
        pmask.append(np.frombuffer(b.raw[:n], np.uint8) if st == 0 else s)
        pstat.append(st)
        poff.append(poff[-1] + n)
    dna, doff, frames, foff, flen = [], [0], [], [0], []
    for n in list(range(0, 25)) + [149, 150, 151, 400]:
        s = np.frombuffer(b"ACGTACGTACGTacgtN", np.uint8)[rng.integers(0, 17, n)].copy()
        out = ctypes.create_string_buffer(2 * n + 16)
        lens = np.zeros(6, np.uint32)
        R.gzr_translate(s.tobytes(), out, oracle_py.ptr(lens))
        t = int(lens.sum())
        dna.append(s)
        doff.append(doff[-1] + n)
        frames.append(np.frombuffer(out.raw[:t], np.uint8))
        foff.append(foff[-1] + t)
        flen.append(lens)
    np.savez_compressed(path, protein=np.concatenate(prot), protein_offsets=np.array(poff),
                        masked=np.concatenate(pmask), seg_status=np.array(pstat),
                        dna=np.concatenate(dna), dna_offsets=np.array(doff),
                        frames=np.concatenate(frames), frame_offsets=np.array(foff),
                        frame_lengths=np.array(flen))


def db_index(path):
    import synthetic
    (_, _, _), (dn, dr, do) = synthetic.protein_workload(20261206, 40, 8, qlen=(100, 200),
                                                         homologs=6)
    db = gz_native.Database(dn, dr, do)
    seq = db.chunk(0)["sequence"]
    m, nl = ref_matrix()
    out = {"sequence": seq}
    for clustering in (1, 0):
        h = R.gzr_build_index(oracle_py.ptr(seq), seq.size, clustering, 10, 39, oracle_py.ptr(m),
                              nl, b"A KR EDNQ C G H ILVM FYW P ST")
        sizes = np.zeros(7, np.uint64)
        R.gzr_index_sizes(h, oracle_py.ptr(sizes))
        arrays = [np.zeros(max(int(s), 1), np.uint32) for s in sizes]
        ptrs = (ctypes.c_void_p * 7)(*[a.ctypes.data for a in arrays])
        R.gzr_index_copy(h, ptrs)
        R.gzr_index_free(h)
        for k, a, s in zip(["hashes", "rep_offsets", "rep_positions", "rep_member_offsets",
                            "member_positions", "value_offsets", "value_positions"], arrays,
                           sizes):
            out[f"c{clustering}_{k}"] = a[:int(s)]
    assert out["c1_member_positions"].size > 0, "no cluster members in the fixture"
    np.savez_compressed(path, **out)


def db_test_protein(directory):
    """The reference test's test_protein.fa (test/test_protein.fa) as a
    database in the reference's file set, chunked at 60 residues as
    test/database_test.cpp:31 does: chunk files in the DatabaseChunk::Save
    layout, .sdp / .scp from the reference's own SeedSearcherDatabaseParameters
    (db defaults: -C T -L 10 -s 39)."""
    import shutil
    os.makedirs(directory, exist_ok=True)
    fa = os.path.join(directory, "test_protein.fa")
    shutil.copyfile(os.path.join(REFERENCE, "test", "test_protein.fa"), fa)
    db = gz_native.Database.from_fasta(fa, chunk_size=60)
    i = db.info()
    chunks = []
    for c in range(i.number_chunks):
        ch = db.chunk(c)
        chunks.append((ch["sequence"], ch["offsets"], ch["names"]))
    m, nl = ref_matrix()
    oracle_py.write_reference_database(os.path.join(directory, "test_protein"), chunks,
                                       i.max_length, i.total_length, i.total_sequences,
                                       True, 10, 39, m, nl)


def main():
    global R
    if not oracle_py.ref_available():
        sys.exit("oracle/_ref is not built (the reference tree is needed once)")
    R = oracle_py.ref()
    R.gzr_read_matrix.restype = ctypes.c_int
    makers = {"gapped_extender.npz": gapped,
              "ungapped_distance_hash.npz": ungapped_distance_hash,
              "statistics.npz": statistics,
              "seg_translate.npz": seg_translate,
              "db_index.npz": db_index,
              "db_test_protein": db_test_protein}
    for name in sys.argv[1:] or list(makers):
        makers[name](os.path.join(HERE, name))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
