import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tf-pwa_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

CONFIG_DIR = os.path.join(ROOT, "tf-pwa_amd", "tf_pwa_amd", "configs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def config_path(name):
    return os.path.join(CONFIG_DIR, name)


@pytest.fixture(scope="session")
def gpu_available():
    from tf_pwa_amd import _lib

    n = __import__("ctypes").c_int(0)
    _lib.check(_lib.lib().tpa_device_count(__import__("ctypes").byref(n)))
    if n.value < 1:
        pytest.fail("no GPU visible to the engine")
    return True
