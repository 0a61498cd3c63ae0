import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests (fails loudly if the library or GPU is missing)."""
    from fast_amd.tau import Context
    c = Context(0)
    yield c
    c.close()
