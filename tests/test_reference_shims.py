"""The class-level drop-in: tests/shim/reference_shims.h gives the
reference's class and method names (DistanceCalculatorGpu,
UngappedExtenderGpu, GappedExtenderGpu, HostSeedsMemory, GpuStreamController,
SeedSearcherGpu) over include/ghostz_hip.h, and tests/shim/shim_driver.cpp
drives them the way AlignerGpuPresearchThread does
(src/aligner_gpu_presearch_thread.cpp:150-410).  The driver is compiled with
g++ against the header alone (the CPU test) and run on the GPU, where every
value it prints is checked against the oracle: the seed searcher's hits per
query in enumeration order, the representative distance pairs and their
distances, ungapped extensions with trigger (passing and failing seeds) and
both sides of the gapped extension of every hit, all through the stream
controller's two resources.
"""
import ctypes
import os
import subprocess
import sys
from ctypes import POINTER, c_uint32, c_void_p

import numpy as np
import pytest

import gz_native as gz
import oracle_py
import synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "shim")
DRIVER = os.path.join(SHIM, "shim_driver")


def test_shims_compile_against_the_header_alone():
    r = subprocess.run(["make", "-C", SHIM, "-B"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert os.path.exists(DRIVER)


def oracle_seed_hits(qv, chunks, info, m, nl, chunk):
    L = oracle_py.oracle()
    L.gzo_seed_hits.restype = ctypes.c_int
    L.gzo_seed_hits.argtypes = [c_void_p, c_void_p, c_void_p, c_uint32, POINTER(c_void_p),
                                POINTER(c_void_p)]
    keep = []
    q, db, arr = oracle_py._structs(qv, chunks, info, keep)
    p = oracle_py._Parameters(arr(m, np.int32), nl, 25, 11, 1, 7.0, 22.0, 15.0, 25.0, 10, 1)
    po, ph = c_void_p(), c_void_p()
    assert L.gzo_seed_hits(ctypes.byref(q), ctypes.byref(db), ctypes.byref(p), chunk,
                           ctypes.byref(po), ctypes.byref(ph)) == 0
    n = int(qv["number_queries"])
    off = np.frombuffer(ctypes.string_at(po, 4 * (n + 1)), np.uint32).copy()
    hits = np.frombuffer(ctypes.string_at(ph, 8 * int(off[-1])), np.uint32).reshape(-1, 2).copy() \
        if off[-1] else np.zeros((0, 2), np.uint32)
    L.gzo_free(po)
    L.gzo_free(ph)
    return off, hits


def parse(path):
    rec = {k: [] for k in "CHDUG"}
    for ln in open(path):
        f = ln.split()
        rec[f[0]].append([int(x) for x in f[1:]])
    return {k: np.array(v, np.int64).reshape(-1, {"C": 2, "H": 4, "D": 5, "U": 4,
                                                  "G": 7}[k]) for k, v in rec.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["protein_two_chunks", "translated"])
def test_shim_driver_matches_oracle(tmp_path, case):
    assert os.path.exists(DRIVER), "build first: make -C tests/shim"
    gz.check(gz.lib().gz_set_device(0))
    if case == "protein_two_chunks":
        (qn, qr, qo), (dn, dr, do) = synthetic.protein_workload(777, 1500, 100)
        translated, chunk = False, int(do[-1]) // 2 + 2048
    else:
        (qn, qr, qo), (dn, dr, do) = synthetic.translated_workload(778, 1200, 150)
        translated, chunk = True, 0
    qf, df, out = tmp_path / "q.fasta", tmp_path / "d.fasta", tmp_path / "out.txt"
    synthetic.write_fasta(str(qf), qn, qr, qo)
    synthetic.write_fasta(str(df), dn, dr, do)
    print("shim driver case", case, flush=True)
    args = [DRIVER, str(qf), str(df), str(int(translated)), str(out)] + ([str(chunk)] if chunk
                                                                          else [])
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    got = parse(str(out))

    q = gz.Queries.from_fasta(str(qf), translated=translated)
    db = gz.Database.from_fasta(str(df), chunk_size=chunk)
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    assert len(got["C"]) == info.number_chunks
    if case == "protein_two_chunks":
        assert info.number_chunks == 2
    qv = q.view()
    m, nl = gz.blosum62()
    offs = qv["offsets"].astype(np.int64)
    nq = int(qv["number_queries"])
    coff = offs[:-1] - np.arange(nq)          # shared-delimiter concatenation
    concat = np.concatenate([qv["codes"][offs[i]:offs[i + 1] - 1] for i in range(nq)] +
                            [np.array([25], np.uint8)])
    cut = np.zeros(nq, np.int32)
    trig = np.zeros(nq, np.int32)
    gcut = ctypes.c_int()
    gz.check(gz.lib().gz_queries_cutoffs(q.handle, None, gz.ptr(cut), gz.ptr(trig),
                                         ctypes.byref(gcut)))
    O = oracle_py.oracle()
    lists = oracle_py.query_lists(qv, chunks, info)
    rmap = np.array(list(info.reduced_code_map), np.uint8)
    total_hits = 0
    for ci in range(info.number_chunks):
        seq = chunks[ci]["sequence"]
        # seed search hits, per query in enumeration order
        off, want = oracle_seed_hits(qv, chunks, info, m, nl, ci)
        H = got["H"][got["H"][:, 0] == ci]
        for i in range(nq):
            mine = H[H[:, 1] == i][:, 2:]
            ref = want[off[i]:off[i + 1]].astype(np.int64)
            assert np.array_equal(mine, ref), (case, ci, i)
        total_hits += len(H)
        # representative distance pairs of the first lists and their distances
        D = got["D"][got["D"][:, 0] == ci]
        ch = chunks[ci]
        for li in sorted(set(D[:, 1].tolist()))[:40]:
            rows = D[D[:, 1] == li]
            h = lists["hash"][li]
            k = np.searchsorted(ch["hashes"], h)
            assert k < len(ch["hashes"]) and ch["hashes"][k] == h
            reps = ch["rep_positions"][ch["rep_offsets"][k]:ch["rep_offsets"][k + 1]]
            e0, e1 = lists["entry_offsets"][li], lists["entry_offsets"][li + 1]
            ent = [(int(coff[lists["entry_queries"][e]] + lists["entry_positions"][e]))
                   for e in range(e0, e1) if lists["entry_checks"][e]]
            pairs = [(a, int(r)) for r in reps for a in ent]
            assert [tuple(x) for x in rows[:, 2:4].tolist()] == pairs, (case, ci, li)
        for _, _, a, b, dist in D[:3000]:
            assert dist == O.gzo_distance(gz.ptr(concat), int(a), gz.ptr(seq), int(b), 10,
                                          gz.ptr(rmap))
        # ungapped extension with trigger
        U = got["U"][got["U"][:, 0] == ci]
        assert len(U) >= len(H)
        qid = np.searchsorted(coff, U[:, 1], side="right") - 1
        flags = np.array([O.gzo_ungapped_with_trigger(gz.ptr(concat), int(a), gz.ptr(seq), int(b),
                                                      25, gz.ptr(m), nl, int(cut[k]),
                                                      int(trig[k])) > trig[k]
                          for (_, a, b, _), k in zip(U, qid)], np.int64)
        assert np.array_equal(U[:, 3], flags), case
        assert 0 < flags.sum() < len(flags)
        # both sides of every hit
        G = got["G"][got["G"][:, 0] == ci]
        assert len(G) == 2 * len(H)
        for _, rev, a, b, s, eq, ed in G:
            ws, p0, p1, _, _ = oracle_py.gapped_one_side("oracle", concat, int(a), concat.size, seq,
                                                          int(b), 25, bool(rev), m, nl, 11, 1,
                                                          gcut.value)
            assert (s, eq, ed) == (ws, int(a) + p0, int(b) + p1), (case, rev, a, b)
    assert total_hits > 50
