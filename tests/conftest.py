import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle_py
    return oracle_py


@pytest.fixture(scope="session")
def ref_lib():
    import oracle_py
    r = oracle_py.ref()
    if r is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    return r
