import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "highordermri.jl_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhoe.so")


@pytest.fixture(scope="session")
def gpu_lib():
    """libhoe loaded on a machine with a visible GPU (GPU tests only)."""
    import hoe
    lib = hoe.load()
    n = lib.hoe_device_count()
    assert n > 0, "no HIP device visible to libhoe"
    return lib
