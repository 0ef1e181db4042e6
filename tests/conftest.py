import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def _has_gpu():
    try:
        import topologyoptimization_amd as T
        n, _ = T.device_info()
        return n > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    import topologyoptimization_amd as T
    n, name = T.device_info()
    if n == 0:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU box")
    print("device:", name, flush=True)
    return T


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
