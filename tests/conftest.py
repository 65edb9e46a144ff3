import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ac-dc_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device (MI355X)")


@pytest.fixture(scope="session")
def acdc():
    import acdc_amd
    return acdc_amd


@pytest.fixture(scope="session")
def orc():
    import oracle_bind
    oracle_bind.lib()
    return oracle_bind
