import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")
    config.addinivalue_line("markers", "reference: reads fixtures from the reference tree at test time")


@pytest.fixture(scope="session")
def orc():
    import oracle

    if not os.path.exists(os.path.join(oracle.HERE, "libqb_oracle.so")):
        oracle.build(ref=True)
    return oracle


def need_reference():
    if not os.path.isdir("/root/reference"):
        pytest.skip("reference tree not present (fixtures are read in place)")
