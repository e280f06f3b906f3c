import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PKG = os.path.join(ROOT, "trotter-suzuki-mpi_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _session_may_run_gpu_tests(config) -> bool:
    expr = " ".join((config.option.markexpr or "").split())
    return "not gpu" not in expr and os.path.exists("/dev/kfd")


def pytest_sessionstart(session):
    """Starts every C++ program the tests need now, before any test has touched
    the GPU in this process (see tests/spawned.py)."""
    config = session.config
    if config.option.collectonly:
        return
    import spawned
    spawned.start(gpu=_session_may_run_gpu_tests(config))


def pytest_sessionfinish(session, exitstatus):
    import spawned
    spawned.finish()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
