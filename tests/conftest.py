import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch's HIP runtime first, before any test loads the library (ctypes or
# pydense), so that both share it (see opensfm_amd._share_torch_runtime).
import torch  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pm_oracle
    pm_oracle.lib()
    return pm_oracle


@pytest.fixture(scope="session")
def pydense():
    """The native module; a GPU test fails (does not skip) when it or the device is missing."""
    from opensfm_amd import native
    mod = native()
    if not mod.DepthmapEstimator.is_available():
        pytest.fail("no HIP device visible to the native library")
    return mod


@pytest.fixture(scope="session")
def scene_small():
    """Three-plane scene, 320x240, four neighbours (seed 1)."""
    from opensfm_amd import synthetic
    return synthetic.three_planes(1, base_cell=0.3, supersample=2)


@pytest.fixture(scope="session")
def scene_tiny():
    """Three-plane scene cut down to 96x72 with two neighbours (fast oracle runs)."""
    from opensfm_amd import synthetic
    return synthetic.three_planes(3, width=96, height=72, num_src=2, base_cell=0.3, supersample=2)
