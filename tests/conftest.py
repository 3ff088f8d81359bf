import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsadmrg.so")


@pytest.fixture(scope="session")
def sad():
    import sadmrg_testlib as tl
    return tl.load_pkg()


@pytest.fixture(scope="session")
def ctx(sad):
    c = sad.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def cgc(sad):
    return sad.Cgc(1)
