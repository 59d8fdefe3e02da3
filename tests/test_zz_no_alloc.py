"""The allocation-free gapped extension entry point
(gz_gapped_extender_extend_one_side_score_only_noalloc_async): its workspace
is a static device array, and its results must equal the allocating entry
point's (score, end positions, DP cells), on two calls in a row with the
second smaller, with bands that outgrow both LDS rings so that the overflow
lists and the global ring in the static array are used.  Sizes past its
stated maximum are refused before anything runs."""
import ctypes
import sys

import numpy as np
import pytest

import gz_native as gz
from helpers import DELIM, related_pair_set

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def extender():
    gz.check(gz.lib().gz_set_device(0))
    m, nl = gz.blosum62()
    rng = np.random.default_rng(4242)
    q, qo, d, do, qpos, dpos = related_pair_set(rng, 3000, min_len=40, max_len=900,
                                                identity=(0.4, 0.98), indel=0.05)
    L = gz.lib()
    _, qp = gz.pack_on_device(q)
    _, dp = gz.pack_on_device(d)
    dm = gz.DeviceBuffer.from_array(m)
    h = ctypes.c_void_p()
    gz.check(L.gz_gapped_extender_create(ctypes.byref(h)))
    gz.check(L.gz_gapped_extender_set_queries(h, DELIM, qp.ptr))
    gz.check(L.gz_gapped_extender_set_database(h, dp.ptr))
    gz.check(L.gz_gapped_extender_set_score_parameters(h, dm.ptr, nl, 11, 1, 52))
    gz.check(L.gz_gapped_extender_set_max_sequence_length(h, 1000))
    yield {"h": h, "qpos": qpos, "dpos": dpos, "keep": (qp, dp, dm)}
    L.gz_gapped_extender_destroy(h)


def run(ext, qpos, dpos, reverse, noalloc):
    L = gz.lib()
    n = qpos.size
    s = gz.Stream()
    a, b = (qpos - 1, dpos - 1) if reverse else (qpos, dpos)
    d_q = gz.DeviceBuffer.from_array(a.astype(np.uint32))
    d_d = gz.DeviceBuffer.from_array(b.astype(np.uint32))
    d_s, d_c = gz.DeviceBuffer(n * 4), gz.DeviceBuffer(n * 4)
    fn = (L.gz_gapped_extender_extend_one_side_score_only_noalloc_async if noalloc else
          L.gz_gapped_extender_extend_one_side_score_only_cells_async)
    gz.check(fn(ext["h"], n, int(reverse), d_q.ptr, d_d.ptr, d_s.ptr, d_c.ptr, s.handle))
    s.sync()
    f = (ctypes.c_uint32 * 1)()
    gz.check(L.gz_gapped_extender_failures(ext["h"], f))
    assert f[0] == 0
    return [x.to_array(t, n) for x, t in ((d_s, np.int32), (d_q, np.uint32), (d_d, np.uint32),
                                           (d_c, np.uint32))]


def test_noalloc_equals_the_allocating_entry_point(extender):
    qpos, dpos = extender["qpos"], extender["dpos"]
    narrow, ring64 = (ctypes.c_uint32 * 1)(), (ctypes.c_uint32 * 1)()
    for n in (qpos.size, qpos.size // 3):
        for reverse in (False, True):
            say("no-alloc gapped extension: jobs", n, "reverse", reverse)
            want = run(extender, qpos[:n], dpos[:n], reverse, noalloc=False)
            gz.check(gz.lib().gz_gapped_extender_overflows(extender["h"], narrow, ring64))
            got = run(extender, qpos[:n], dpos[:n], reverse, noalloc=True)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
    # the inputs send bands through both overflow lists and the global ring
    assert narrow[0] > 0 and ring64[0] > 0


def test_noalloc_refuses_sizes_past_its_maximum(extender):
    L = gz.lib()
    st = L.gz_gapped_extender_extend_one_side_score_only_noalloc_async(
        extender["h"], (1 << 20) + 1, 0, None, None, None, None, None)
    assert st == 5 and b"1048576" in L.gz_last_error()
    gz.check(L.gz_gapped_extender_set_max_sequence_length(extender["h"], 5000))
    st = L.gz_gapped_extender_extend_one_side_score_only_noalloc_async(
        extender["h"], 10, 0, None, None, None, None, None)
    gz.check(L.gz_gapped_extender_set_max_sequence_length(extender["h"], 1000))
    assert st == 5 and b"4094" in L.gz_last_error()
