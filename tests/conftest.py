import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vapoursynth-nssfactory_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible device; the test fails (never falls back)
    when the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nss_hip
    nss_hip.lib()
    return nss_hip
