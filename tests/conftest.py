import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def have_gpu():
    """True when a HIP device is visible, without initialising HIP in this
    process (tests that need one load the library themselves)."""
    return os.path.exists("/dev/kfd") and any(
        n.startswith("renderD") for n in os.listdir("/dev/dri")) if os.path.isdir("/dev/dri") else False


@pytest.fixture(scope="session")
def built():
    """Paths of the built product; fails loudly when it is not built."""
    import spatialjoincountovershells_amd as sj
    for p in (sj.LIB_PATH, sj.HOST_LIB_PATH, sj.CLI_PATH):
        if not os.path.exists(p):
            pytest.fail(f"{p} is not built: python -c 'import __graft_entry__ as g; g.build()'")
    return sj
