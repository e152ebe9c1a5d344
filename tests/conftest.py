import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "din-mol-li_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU checks")
    if os.environ.get("DML_SIMT"):
        # the GPU tests against the CPU SIMT build of the same sources
        # (tests/simt: fibers per lane, AddressSanitizer); a debugging aid,
        # never the default
        sys.path.insert(0, os.path.join(ROOT, "tests", "simt"))
        import sim
        sim.use()


@pytest.fixture(scope="session")
def orc():
    import orc as _orc
    _orc.build()
    return _orc


@pytest.fixture(scope="session")
def dml():
    """the ctypes binding of libdml (raises if the library is not built)"""
    import dml as _dml
    _dml.lib()
    return _dml
