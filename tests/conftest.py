import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("nematic-worms_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def nwlib():
    """libnwamd.so with a usable device; errors are returned, not fatal."""
    from nwamd import abi
    lib = abi.load()
    lib.nwSetErrorMode(abi.NW_ERRORS_RETURN)
    if not lib.nwDeviceAvailable():
        pytest.fail("no HIP device: GPU tests need an MI355X")
    lib.cudaInit(0, None)
    return lib


@pytest.fixture(autouse=True)
def _gpu_errors_clean(request):
    yield
    if "gpu" in request.keywords and "nwlib" in request.fixturenames:
        lib = request.getfixturevalue("nwlib")
        err = lib.nwGetLastError()
        msg = lib.nwGetLastErrorString()
        assert err == 0, f"library error {err}: {msg}"
