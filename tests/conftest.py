import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_runtest_setup(item):
    if "gpu" in item.keywords:
        print(f"\n[gpu-test] {item.nodeid}", flush=True)


@pytest.fixture(scope="session")
def gpu_lib():
    """The real gfx950 library; GPU tests fail loudly if it is missing."""
    from shafft_amd import api

    L = api.lib()
    api.use(None)
    return L


@pytest.fixture(scope="session")
def emu_lib():
    """CPU emulation build of the host plan code (tests/emu)."""
    import subprocess

    from shafft_amd import api

    path = os.path.join(ROOT, "_build", "emu", "libshafft_emu.so")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "emu")], check=True)
    return api.bind(path)
