import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

EMU_LIB = os.path.join(TESTS, "libgammahip_emu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def emu_lib():
    """CPU emulation of the kernel bodies (built on demand)."""
    if not os.path.exists(EMU_LIB):
        from gamma_amd import build
        build.build_emu()
    return EMU_LIB


def l1rel(a, b):
    import numpy as np
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    assert a.shape == b.shape, (a.shape, b.shape)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))
