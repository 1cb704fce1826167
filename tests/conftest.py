import os
import sys

import pytest

# PyTorch ships its own HIP runtime with the same soname
# (libamdhip64.so.7) as the one the library links; whichever is loaded
# first serves the whole process, and PyTorch finds no GPU on a runtime
# other than its own.  Loading it before any test opens the library lets
# the ensemble, simulation and autograd tests share one process.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (HIP device); parity tests proper")


def gpu_available():
    try:
        import ctypes
        lib = ctypes.CDLL(os.path.join(ROOT, "synaptogenesis_amd",
                                       "libsg_hip.so"))
        return lib.sg_hip_device_count() > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def lib():
    from synaptogenesis_amd import syngen
    L = syngen.load_library()
    L.set_suppress_output(True)
    L.set_warnings(False)
    return L


@pytest.fixture(scope="session")
def gpu_lib(lib):
    if not gpu_available():
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return lib
