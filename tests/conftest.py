import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "vlite-frb_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def vfrb():
    import vfrb as m
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle as m
    return m


@pytest.fixture(scope="session")
def gpu(vfrb):
    """Device fixture for GPU tests: fails (never skips) without a GPU."""
    n = vfrb.device_count()
    assert n > 0, "no GPU visible to libvfrb"
    print(f"[gpu] {vfrb.device_name()}", flush=True)
    return vfrb
