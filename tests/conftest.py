import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


def gpu_available() -> bool:
    try:
        from gpuvmem_amd import native
        import ctypes as C
        lib = native.load_hip()
        n = C.c_int()
        return lib.gvm_device_info(None, 0, C.byref(n), None) == 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    from gpuvmem_amd import native
    return native.load_hip()
