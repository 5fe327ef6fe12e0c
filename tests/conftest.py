import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP device)")
    # Build the checker and the library when they are missing (CPU container);
    # on a GPU box both arrive prebuilt with the tree.
    from oracle import bind
    if not os.path.exists(bind.LIB_PATH):
        bind.build()
    lib = os.path.join(ROOT, "simulator_amd", "libclothcore.so")
    if not os.path.exists(lib):
        from simulator_amd import build
        build.build()


def gpu_available():
    from simulator_amd import _lib
    return _lib.device_count() > 0


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True


@pytest.fixture(params=["host", pytest.param("gpu box", marks=pytest.mark.gpu)])
def both_suites(request):
    """A module that uses this fixture runs each of its tests twice: in the
    CPU suite and again in the GPU suite, so that the parity of host-side
    rows (spring builders, OBJ reader, boundary springs) is part of the GPU
    box's run too."""
    return request.param
