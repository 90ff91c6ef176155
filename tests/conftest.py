import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def renderer():
    """One HIP context for the whole GPU session (fails loudly without a GPU)."""
    from blackhole_8_amd import Renderer

    r = Renderer(0)
    print(f"[gpu] device: {r.device_info()}", flush=True)
    yield r
    r.close()
