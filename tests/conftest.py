import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (run with -m gpu)")


def pytest_collection_modifyitems(config, items):
    # GPU tests need the built HIP library to load (a missing library is an
    # error, not a skip); they are skipped only when no device is present.
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    from ghostlight_amd import _abi
    lib = _abi.load_library()
    if not lib.gl_device_available():
        skip = pytest.mark.skip(reason="no gfx950 device in this environment")
        for it in gpu_items:
            it.add_marker(skip)
