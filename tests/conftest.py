import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "hnsw-index_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP kernels)")


def pytest_collection_modifyitems(config, items):
    # GPU tests need a device; without one they are skipped, never faked.
    import ctypes
    lib_path = os.path.join(ROOT, "hnsw-index_amd", "libhnsw_amd.so")
    have = False
    if os.path.exists(lib_path):
        try:
            have = ctypes.CDLL(lib_path).hh_device_count() > 0
        except OSError:
            have = False
    if have:
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
