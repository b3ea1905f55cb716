import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "btc-tc-artifact_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def btc():
    import btc_amd
    btc_amd.lib()  # raises if libbtc_amd.so was not built: no silent fallback
    return btc_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle._lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(btc):
    """The HIP device; GPU tests fail loudly when the library is missing and
    are skipped only where no device is visible (the CPU container)."""
    rc = btc.lib().btc_device_init(0)
    if rc != 0:
        pytest.skip("no HIP device: " + btc.lib().btc_last_error().decode())
    return btc
