import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")
    config.addinivalue_line("markers", "sanitize: runs the CPU suite or a harness under a sanitizer build")


@pytest.fixture(scope="session")
def ctx():
    from sigmo_amd import capi

    # No explicit close: device objects made in module fixtures hold a reference to the
    # context and may be collected after session teardown; the context is destroyed when the
    # last of them is.
    yield capi.Context(0)
