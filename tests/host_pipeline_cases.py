"""Runs the host-flow cases of test_host_pipeline.py (all, or those named
after the library path) against a given libhostpipe build (a child process
of test_host_sanitizer.py, which loads a sanitizer build).  Exits non-zero
on a mismatch."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ghostz-gpu_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import oracle_py  # noqa: E402

oracle_py.HOSTPIPE_PATH = sys.argv[1]
import test_host_pipeline as T  # noqa: E402

only = set(sys.argv[2:])
for name in sorted(dir(T)):
    if name.startswith("test_") and (not only or name in only):
        fn = getattr(T, name)
        if name == "test_thread_count_does_not_change_output":
            for t in (1, 3):
                fn(t)
        else:
            fn()
        print("ok", name, flush=True)
