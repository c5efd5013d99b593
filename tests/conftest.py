import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import bindings
    if not os.path.exists(bindings.ORACLE_SO):
        bindings.build(ref=False)
    return bindings.Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import bindings
    if not bindings.ref_available():
        pytest.skip("reference engine build (oracle/_ref) not available")
    return bindings.Reference()


@pytest.fixture(scope="session")
def gpu():
    import flipengine3d_amd as fa
    return fa.Context(0)
