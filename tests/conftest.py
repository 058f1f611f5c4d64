import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mini-fem-pic-main_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_EXAMPLES = "/root/reference/examples"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C ABI")


@pytest.fixture(scope="session")
def mp():
    import minipic_amd

    minipic_amd.lib()
    return minipic_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle
