import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
NI = os.path.join(GOLDEN, "Ni_master_20kV.npz")
GAN = os.path.join(GOLDEN, "GaN_master_20kV.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libzebsd_hip.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ni_master():
    from zebsdsolve_amd import master
    return master.load(NI)


@pytest.fixture(scope="session")
def gan_master():
    from zebsdsolve_amd import master
    return master.load(GAN)
