import os
import sys

import pytest

# libhh.so and PyTorch each bring a HIP runtime of the same soname, and a
# process keeps the copy it loads first; PyTorch only initialises its
# devices on its own.  Tests that capture graphs through PyTorch therefore
# import it before any test loads libhh.so.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        return torch is not None and torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def xavier_w():
    from hignn_amd import weights
    return weights.xavier(0)


@pytest.fixture(autouse=True, scope="session")
def _batched_far_field_by_default():
    """The library picks the persistent far-field kernel for systems with few
    far blocks; the GPU tests cover the batched path unless a test sets
    HH_FAR_MODE itself (test_far_mode_auto tests the automatic choice)."""
    os.environ.setdefault("HH_FAR_MODE", "1")
    yield
