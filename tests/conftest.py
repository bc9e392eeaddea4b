import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# tests compile every pattern they construct: a table cache directory from the
# environment would let one test's compiled tables stand in for another's
os.environ.pop("TREX_AMD_TABLE_CACHE", None)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def gpu_lib():
    """The built library on a live gfx950 device; a GPU test fails (never skips)
    when either is missing, so a passing GPU run always exercised the HIP path."""
    from trex_amd import _lib
    lib = _lib.lib()
    rc = lib.trex_device_check()
    assert rc == 0, f"trex_device_check() = {rc}: {_lib.last_error()}"
    return lib
