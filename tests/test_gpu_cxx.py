"""The C++ drop-in API (mrpnn::VolumeRender / mrpnn::Camera) end to end on the GPU: the
reference's README flow renders predictor and path-traced PNGs."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
DEMO = os.path.join(ROOT, "tests", "cpp", "build", "cxx_api_demo")


def test_cxx_demo_renders_pngs(tmp_path):
    if not os.path.exists(DEMO):
        pytest.skip("demo not built (run __graft_entry__.build())")
    r = subprocess.run([DEMO, str(tmp_path), "48"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    active = int(words[words.index("active") + 1])
    mean_pred = float(words[words.index("mean_pred") + 1])
    mean_pt = float(words[words.index("mean_pt") + 1])
    assert active > 100 and mean_pred > 0 and mean_pt > 0
    # interactive frames (PT with the filter, then MRPNN) light part of the 64x64 RGBA8 image
    # (the soft ball covers about 200 pixels of the 64x64 view)
    assert int(words[words.index("lit_pt") + 1]) > 50 and int(words[words.index("lit_pred") + 1]) > 50, r.stdout
    # RPNN frame of the same camera (its own sub-pixel jitter): about the same active pixels, lit
    active_rpnn = int(words[words.index("active_rpnn") + 1])
    assert abs(active_rpnn - active) <= 0.05 * active and float(words[words.index("mean_rpnn") + 1]) > 0
    for name in ("demo_mrpnn.png", "demo_rpnn.png", "demo_pt.png"):
        data = open(tmp_path / name, "rb").read()
        assert data[:8] == b"\x89PNG\r\n\x1a\n" and len(data) > 96 * 96 * 4
