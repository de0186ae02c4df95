import gc
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "highvoronoi.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhv.so)")


@pytest.fixture(scope="module", autouse=True)
def _release_module_contexts():
    # a module's device contexts (and their memory pools) go with the module, also where a
    # reference cycle holds one, before the next module allocates its own; so do the contexts
    # a batched build keeps for the process
    yield
    gc.collect()
    lib = sys.modules.get("hvgpu._lib")
    if lib is not None and getattr(lib, "_lib", None) is not None:
        lib.release_batch(-1)
