import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="session", autouse=True)
def _strict_extra_memory():
    """Extra device memory is checked against alpha itself: no holding floor
    (test_extra_memory_follows_alpha covers the default floor)."""
    try:
        import eithot_amd as E
        old = E.set_tuning("hold_floor", 0)
    except Exception:  # the library is not built: nothing to configure
        yield
        return
    yield
    E.set_tuning("hold_floor", old)
