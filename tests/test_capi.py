"""The C ABI boundary without a GPU: the library loads, exports every function declared in
include/mrp_c_api.h, and rejects bad arguments with status codes (no compute calls)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "mrp_c_api.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mrp_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    import mrpnn_amd as M
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("library not built")
    return M.lib()


def test_exports_every_declared_symbol(lib):
    names = declared()
    assert len(names) >= 30
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "mrpnn_amd", "libmrpnn_hip.so")],
                        capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)", nm))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_argument_errors_are_status_codes(lib):
    import mrpnn_amd as M
    assert lib.mrp_ctx_create(None, None) == 1  # MRP_ERR_INVALID
    assert b"null" in lib.mrp_last_error()
    assert lib.mrp_volume_create(None, 64, None) == 1
    h = ctypes.c_void_p()
    assert lib.mrp_volume_create(None, 64, ctypes.byref(h)) == 1
    assert lib.mrp_render_plain(None, None, None, 1, None) == 1
    assert lib.mrp_render_batch(None, None, 0, None, 0, M.BF16, None) == 1
    assert lib.mrp_render_batch_noalloc(None, None, 0, None, 0, M.BF16, None) == 1
    assert lib.mrp_pt_render(None, None, None, 0, 1, 1, 0, 1.0, None, None) == 1
    assert lib.mrp_predict_features(None, None, 1, None, 0.5, 0, None) == 1
    # grid + density mips 0..8 + TR mips 0..7 + zero-padded copies ((res+3)^3) of density
    # mips 0..7 + (y, z) quads (4 floats per padded texel) of the padded TR mips 0..6
    pad_d = sum(((256 >> m) + 3) ** 3 for m in range(8))
    pad_t = 4 * sum(((128 >> m) + 3) ** 3 for m in range(7))
    # + (y, z) quads (16 bytes per padded texel) of density mips 0..7 (fast gather, TR bake)
    quads = sum(((256 >> m) + 3) ** 3 for m in range(8))
    # + one emptiness byte per 8^3 brick of the march ((N + 8) // 8 per axis)
    assert lib.mrp_volume_bytes(512) == (512 ** 3 + 19173961 + 2396745 + pad_d + pad_t) * 4 + quads * 16 + 65 ** 3


def test_version_string(lib):
    assert b"gfx950" in lib.mrp_version()


def test_cxx_api_compiles_and_links(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = os.path.join(ROOT, "tests", "cpp", "cxx_api_demo.cpp")
    exe = str(tmp_path / "demo")
    subprocess.check_call([hipcc, "-O1", "-std=c++17", src, "-I" + os.path.join(ROOT, "include"),
                           "-L" + os.path.join(ROOT, "mrpnn_amd"), "-lmrpnn_hip",
                           "-Wl,-rpath," + os.path.join(ROOT, "mrpnn_amd"), "-o", exe])
    assert os.path.exists(exe)
