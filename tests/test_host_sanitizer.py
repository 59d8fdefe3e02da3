"""The product's host flow (query lists, chain filter, AddResults,
BuildResults, output; device stages emulated) under AddressSanitizer and
ThreadSanitizer, on the cases of test_host_pipeline.py.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "oracle", "libhostpipe_asan.so")
TSAN_LIB = os.path.join(ROOT, "oracle", "libhostpipe_tsan.so")


def _runtime(name):
    r = subprocess.run(["g++", f"-print-file-name={name}"], stdout=subprocess.PIPE, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def _libasan():
    return _runtime("libasan.so")


@pytest.mark.skipif(not os.path.exists(ASAN_LIB) or _libasan() is None,
                    reason="AddressSanitizer build or runtime not available")
def test_host_flow_is_clean_under_address_sanitizer():
    env = dict(os.environ, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    # every case but the two largest (the small many-block case covers them)
    cases = ["test_protein", "test_translated", "test_unclustered_and_no_filter",
             "test_query_and_database_chunks", "test_result_limits",
             "test_thread_count_does_not_change_output", "test_edge_queries",
             "test_small_many_blocks"]
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_pipeline_cases.py"),
                        ASAN_LIB] + cases, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:]


@pytest.mark.skipif(not os.path.exists(ASAN_LIB) or _libasan() is None,
                    reason="AddressSanitizer build or runtime not available")
def test_seed_word_walk_reads_stay_inside_the_packed_arrays():
    # the pair walk's read-ahead on exactly sized arrays (ungapped.h)
    env = dict(os.environ, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "test_seed_words.py"),
                        ASAN_LIB], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    assert r.returncode == 0 and "bounds cases done" in r.stdout, r.stdout[-4000:]


@pytest.mark.skipif(not os.path.exists(TSAN_LIB) or _runtime("libtsan.so") is None,
                    reason="ThreadSanitizer build or runtime not available")
def test_batched_pipeline_is_clean_under_thread_sanitizer():
    # the batched form's seeder, extender, AddResults and writer threads
    env = dict(os.environ, LD_PRELOAD=_runtime("libtsan.so"), TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_pipeline_cases.py"),
                        TSAN_LIB, "test_small_many_blocks", "test_small_blocks_two_extenders"],
                       env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ThreadSanitizer" not in r.stdout, r.stdout[-4000:]
