import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libmrpnn_hip.so")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (test infrastructure), built on demand where a compiler exists."""
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mrpnn_amd as M
    c = M.Context()
    yield c
    c.close()


@pytest.fixture(scope="session")
def small_scene(orc):
    """64^3 seeded fBm cloud (5 octaves, base frequency 4) with its oracle pyramids."""
    N = 64
    grid = orc.fbm_volume(N, 5, 4.0, 1)
    sc = orc.Scene(grid, N)
    L = np.array([0.866, 0.5, 0.0], np.float32)
    L /= np.linalg.norm(L)
    sc.update_tr(L, 64.0)
    return sc, L


def have_reference():
    return os.path.isdir(os.path.join(REFERENCE, "core"))
