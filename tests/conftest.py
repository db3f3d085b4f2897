import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vhs-decode-dotnet_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def sinc_lut():
    import numpy as np
    return np.load(os.path.join(PKG, "vhsdecode_hip", "data", "sinc_lut.npz"))["downscale_sinc_lut"]
