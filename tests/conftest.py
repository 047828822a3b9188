import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhtool_amd kernels)")
    config.addinivalue_line("markers", "ref: needs the reference driver oracle/_ref/ref_driver")


@pytest.fixture(scope="session")
def orc():
    from oracle import lib
    lib.lib()
    return lib


@pytest.fixture(scope="session")
def ref():
    from oracle import refcase
    if not refcase.available():
        pytest.skip("reference driver not built (needs /root/reference at build time)")
    return refcase


@pytest.fixture(scope="session")
def htg():
    import htool_amd
    from htool_amd import capi
    capi.lib()  # fails loudly when the library is missing
    return htool_amd
