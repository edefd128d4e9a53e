"""Dispatch order of the level-0 TR bake's and the march's tiles (mrpnn_amd/csrc/bake_order.h), on
the host: a bijection from workgroup ranks onto the tile grid in rings of growing distance from the
centre tile, for every grid shape and centre the bake can see and for large march grids."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_ring_order_is_a_bijection_in_ring_order(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "bake_order_check")
    subprocess.check_call([gxx, "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "bake_order_check.cpp"),
                           "-I" + os.path.join(ROOT, "mrpnn_amd", "csrc"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
