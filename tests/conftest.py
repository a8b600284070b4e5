import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

# Under pytest-xdist each worker process runs the oracle's OpenMP loops:
# one thread each, and no spinning, so the workers do not oversubscribe the
# cores (set before the oracle library starts its thread pool).
if "PYTEST_XDIST_WORKER" in os.environ:
    os.environ.setdefault("VR_ORACLE_THREADS", "1")
    os.environ.setdefault("OMP_WAIT_POLICY", "passive")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session", autouse=True)
def built_libraries():
    """Build the oracle and the product library when nothing is built yet."""
    from oracle import LIBPATH as ORACLE_LIB
    from smallproject_amd._lib import LIBPATH as PRODUCT_LIB
    if not ORACLE_LIB.exists():
        from oracle.build import build as build_oracle
        build_oracle()
    if not PRODUCT_LIB.exists():
        from smallproject_amd.build import build as build_product
        build_product()
    yield


@pytest.fixture(scope="session")
def gpu():
    """GPU tests run only where the ROCm device node exists; there, a device
    that cannot be opened is a failure, not a skip."""
    if not os.path.exists("/dev/kfd"):
        pytest.skip("no GPU device node (/dev/kfd)")
    from smallproject_amd import Renderer, default_config
    r = Renderer(default_config(16, 16))
    r.close()
    yield
