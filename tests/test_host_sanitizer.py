"""The product's host flow (query lists, chain filter, AddResults,
BuildResults, output; device stages emulated) under AddressSanitizer, on the
cases of test_host_pipeline.py.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "oracle", "libhostpipe_asan.so")


def _libasan():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(not os.path.exists(ASAN_LIB) or _libasan() is None,
                    reason="AddressSanitizer build or runtime not available")
def test_host_flow_is_clean_under_address_sanitizer():
    env = dict(os.environ, LD_PRELOAD=_libasan(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host_pipeline_cases.py"),
                        ASAN_LIB], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:]
