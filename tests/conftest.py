import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "fe-demo-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def _gpu_available():
    try:
        import fehip

        return fehip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests call through libfemhip; they fail (not skip) when it cannot run."""
    import fehip

    n = fehip.device_count()
    assert n > 0, "no HIP device visible"
    info = fehip.device_info(0)
    print("device 0:", info.name.decode(), info.arch.decode(), "CUs", info.cu_count, flush=True)
    return info
