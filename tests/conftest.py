import os
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmh.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    if not pyoracle.LIB_PATH.exists():
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def mh():
    from rfmanta_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail("libmh.so is not built (python -m rfmanta_amd.build)")
    return _lib.load()


def tie_classes(p_ref, t_ref, p_gpu, t_gpu):
    """Splits prim-id mismatches into the reference's tie classes.

    H1: |dt| <= 1e-6 * t (exact ties, e.g. shared edges);
    H2: |dt| < 1e-4 absolute (the Kensler-Shirley T_EPSILON acceptance band, where the
        winner depends on visit order); anything else is a real mismatch.
    """
    mis = p_ref != p_gpu
    dt = np.abs(t_ref.astype(np.float64) - t_gpu.astype(np.float64))
    scale = np.maximum(np.abs(t_ref.astype(np.float64)), 1e-30)
    h1 = mis & (dt <= 1e-6 * scale)
    h2 = mis & ~h1 & (dt < 1e-4)
    real = mis & ~h1 & ~h2
    return int(h1.sum()), int(h2.sum()), int(real.sum())


def gpu_available() -> bool:
    return os.path.exists("/dev/kfd")
