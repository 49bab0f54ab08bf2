import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def H():
    import hacapk_hmvm_amd
    return hacapk_hmvm_amd


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(H):
    if H.device_count() < 1:
        pytest.skip("no HIP device")
    return True


def reference_path(*parts):
    """Path under the read-only reference checkout, or None when it is absent
    (machines that run the GPU tests do not have it)."""
    p = os.path.join(REFERENCE, *parts)
    return p if os.path.exists(p) else None
