import os
import sys

import pytest

try:
    # PyTorch-ROCm ships its own HIP runtime (libamdhip64.so.7).  Loading it
    # before libbioshell_gpu.so makes the library bind to that same runtime, so
    # one process never holds two HIP runtimes.
    import torch  # noqa: F401
except Exception:  # torch is only needed by the device-pointer tests
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _ensure_built():
    """Build the product and oracle libraries only if nothing is built yet."""
    from bioshell_amd import _native

    if not os.path.exists(_native.LIB_PATH):
        from bioshell_amd import build as b

        b.build()
    from oracle import oracle as O

    if not os.path.exists(O._LIB_PATH):
        O.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_ctx():
    import bioshell_amd as B

    if B.device_count() == 0:
        pytest.fail("no GPU visible: gpu-marked tests need an MI355X")
    ctx = B.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def golden_2jqb():
    import numpy as np

    return np.load(os.path.join(GOLDEN, "2jqb_ensembles.npz"))


@pytest.fixture(scope="session")
def golden_config0():
    import numpy as np

    return np.load(os.path.join(GOLDEN, "config0.npz"))
