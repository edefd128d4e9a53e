"""The C++ drop-in API (mrpnn::VolumeRender / mrpnn::Camera) end to end on the GPU: the
reference's README flow renders predictor and path-traced PNGs; Camera::RenderToFile(..., MRPNN)
equals mrp_render_plain of the same view bit for bit; Camera::GatherData into a device buffer
(main.cu:30-38) equals mrp_gather_tiles bit for bit (two pixel offsets); Camera::Gamma /
Camera::ACES agree with the oracle's display-path tone maps."""
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
    assert int(words[words.index("render_equal") + 1]) == 1 and int(words[words.index("gather_equal") + 1]) == 1
    assert int(words[words.index("gather_active") + 1]) > 20  # the 12x12 camera sees the ball
    import numpy as np
    import oracle as O
    tones = np.array([float(x) for x in r.stdout.split("tones")[1].split("\n")[0].split()]).reshape(3, 2, 3)
    cols = np.array([[0.5, 0.2, 0.1], [2.0, 1.0, 0.25], [0.01, 0.03, 0.02]], np.float32)
    _, ga = O.tonemap(cols, 1.0, 1)
    _, ac = O.tonemap(cols, 1.0, 2)
    np.testing.assert_allclose(tones[:, 0], ga, rtol=2e-6)
    np.testing.assert_allclose(tones[:, 1], ac, rtol=2e-6)
    for name in ("demo_mrpnn.png", "demo_rpnn.png", "demo_pt.png", "demo_camMRPNN.png"):
        data = open(tmp_path / name, "rb").read()
        assert data[:8] == b"\x89PNG\r\n\x1a\n" and len(data) > 96 * 96 * 4
