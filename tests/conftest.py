import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def gpu_available():
    from openlb_amd import abi
    try:
        n = abi.device_count()
    except Exception as exc:  # library or runtime missing: the GPU tests must fail loudly
        pytest.fail(f"GPU backend unavailable: {exc}")
    if n < 1:
        pytest.fail("no HIP device visible")
    return True
