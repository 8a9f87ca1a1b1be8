import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qlassical-quantum-hpc-simulator_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_present() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not gpu_present():
        pytest.skip("no GPU")
    # a GPU is present: the HIP library must load and see it (no silent skip)
    import qlx
    qlx.lib()
    info = qlx.device_info(0)
    assert info["num_cus"] > 0, info
    return qlx
