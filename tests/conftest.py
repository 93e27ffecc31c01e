import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
os.environ.setdefault("MKL_THREADING_LAYER", "GNU")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ref_available():
    from oracle import ref
    return ref.available()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference CHOLMOD (oracle/_ref/libssref.so) not built")
    ref.set_threads(min(8, os.cpu_count() or 1))
    return ref


@pytest.fixture(scope="session")
def gpu():
    import torch  # noqa: F401  (device memory for device-resident cases)
    from suitesparse_amd import cholmod
    cm = cholmod.Common()
    if not cm.start():
        pytest.fail(f"HIP device init failed: status {cm.status}")
    return cm
