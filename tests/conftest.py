import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ghostz-gpu_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref built from the reference sources")


@pytest.fixture(autouse=True)
def _no_torch_in_gpu_tests(request):
    # torch bundles its own HIP runtime; loaded first, it would serve the
    # product library's HIP calls instead of the ROCm the library was built for
    if request.node.get_closest_marker("gpu") is not None:
        assert "torch" not in sys.modules, "a GPU test process imported torch"
    yield


def pytest_collection_modifyitems(config, items):
    import oracle_py
    skip_ref = pytest.mark.skip(reason="oracle/_ref not built (reference tree absent)")
    for it in items:
        if "ref" in it.keywords and not oracle_py.ref_available():
            it.add_marker(skip_ref)
