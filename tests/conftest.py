import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from mgkmc_amd import _lib
    _lib.load()
    return True


@pytest.fixture(autouse=True)
def _run_in_scratch_dir(tmp_path_factory, monkeypatch):
    """The simulators write their run logs under output_dir ("output",
    relative to the working directory); each test runs in a scratch directory
    of its own so that nothing lands in the repository."""
    monkeypatch.chdir(tmp_path_factory.mktemp("cwd"))
