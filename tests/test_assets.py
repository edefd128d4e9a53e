"""Weight asset, stencil tables and the compile-time weight layout."""
import hashlib
import json
import os
import re

import numpy as np
import pytest

from conftest import REFERENCE, ROOT, have_reference

ASSETS = os.path.join(ROOT, "mrpnn_amd", "assets")
CSRC = os.path.join(ROOT, "mrpnn_amd", "csrc")


def test_weight_asset_manifest():
    man = json.load(open(os.path.join(ASSETS, "mrpnn_weights.json")))
    blob = open(os.path.join(ASSETS, "mrpnn_weights.bin"), "rb").read()
    assert hashlib.sha256(blob).hexdigest() == man["sha256"]
    assert len(blob) == 4 * man["total"] == 4 * 49673
    assert len(man["arrays"]) == 146
    off = 0
    for a in man["arrays"]:
        assert a["offset"] == off
        off += a["count"]
    w = np.frombuffer(blob, "<f4")
    assert np.isfinite(w).all()


def test_net_layout_header_matches_manifest():
    man = json.load(open(os.path.join(ASSETS, "mrpnn_weights.json")))
    hdr = open(os.path.join(CSRC, "net_layout.h")).read()
    offs = dict(re.findall(r"constexpr int (\w+) = (\d+);", hdr))
    for a in man["arrays"]:
        assert int(offs[a["name"]]) == a["offset"]


def test_layer_shapes_from_counts():
    # the shapes RadiancePredictInternal (radiancePredict.cu:301-422) implies for each array
    man = {a["name"]: a["count"] for a in json.load(open(os.path.join(ASSETS, "mrpnn_weights.json")))["arrays"]}
    for br in ("L", "L_Tr", "L_Hg"):
        for k, n in zip(range(9), (8, 8, 16, 16, 16, 32, 32, 32, 32)):
            assert man["%s%d1W" % (br, k)] == n * n and man["%s%d1B" % (br, k)] == n
    for br in ("LD", "LD_Tr", "LD_Hg"):
        for k in range(5):
            assert man["%s%d1W" % (br, k)] == 64 and man["%s%d1B" % (br, k)] == 8
    assert man["LSEFin1W"] == 16 * 75 and man["LSEFin2W"] == 6 * 16 and man["LGGSW"] == 8 * 3
    assert man["LC0W"] == 128 * 128 and man["LC1W"] == 64 * 128 and man["LC2W"] == 32 * 64 and man["LXW"] == 16 * 32
    assert man["LX12W"] == 16 and man["LX12B"] == 1


@pytest.mark.skipif(not have_reference(), reason="reference sources not present")
def test_weight_asset_equals_reference_text():
    import importlib.util
    spec = importlib.util.spec_from_file_location("cw", os.path.join(ASSETS, "convert_weights.py"))
    cw = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cw)
    arrays = cw.parse(open(os.path.join(REFERENCE, "core", "NNWeight.cuh")).read())
    w = np.fromfile(os.path.join(ASSETS, "mrpnn_weights.bin"), "<f4")
    ref = np.concatenate([np.asarray(v, np.float32) for _, v in arrays])
    assert np.array_equal(w, ref)


@pytest.mark.skipif(not have_reference(), reason="reference sources not present")
def test_stencil_tables_equal_reference_text():
    import importlib.util
    spec = importlib.util.spec_from_file_location("et", os.path.join(ASSETS, "extract_tables.py"))
    et = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(et)
    st = et.parse_stencil(open(os.path.join(REFERENCE, "core", "vector.cu")).read())
    hdr = open(os.path.join(CSRC, "ref_tables.h")).read()
    body = hdr[hdr.index("kStencilDirs"):hdr.index("};")]
    got = np.array([[float(x) for x in t] for t in re.findall(r"\{([-\d.e]+)f, ([-\d.e]+)f, ([-\d.e]+)f\}", body)],
                   np.float32)
    ref = np.concatenate([np.asarray(st[n], np.float32) for n in et.NAMES])
    assert got.shape == (160, 3) and np.array_equal(got, ref)


def test_stencil_schedule(orc):
    # GetSamples23_ (vector.cu:304-405): rings, mips, offsets (x256)
    want = [(0, 8, 0, 1.0), (8, 16, 1, 2.5 * 0.92), (16, 32, 2, 5.5 * 0.84), (32, 48, 3, 11.5 * 0.76),
            (48, 64, 4, 23.5 * 0.68), (64, 96, 5, 47.5 * 0.6), (96, 128, 6, 95.5 * 0.6), (128, 160, 7, 191.5 * 0.6)]
    for a, b, mip, off in want:
        for i in range(a, b):
            o, m, t, loc = orc.stencil(i)
            assert m == mip and t == 5 and abs(o * 256 - off) < 1e-4
            assert (loc == 0) == (i == 0)
    light = [(160, 0, lambda k: k + 1), (168, 1, lambda k: 9.5 + 2 * k), (176, 2, lambda k: 26.5 + 4 * k),
             (184, 3, lambda k: 60.5 + 8 * k)]
    for a, mip, f in light:
        for k in range(8):
            o, m, t, loc = orc.stencil(a + k)
            assert m == mip and t == 1 and o * 256 == pytest.approx(f(k), abs=1e-5)
