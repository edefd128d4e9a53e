"""Philox4x32-10 (mrpnn_amd/csrc/philox.h, shared by the library and the oracle) against
the Random123 known-answer vectors and against AMD's open-source rocRAND header."""
import os
import shutil
import subprocess

import numpy as np
import pytest

# Random123 kat_vectors, philox4x32_10: (counter, key) -> output
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(orc, ctr, key, want):
    assert tuple(int(x) for x in orc.philox(ctr, key)) == want


def test_philox_matches_rocrand(orc, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = os.path.join(os.path.dirname(__file__), "cpp", "rocrand_philox_kat.hip")
    exe = str(tmp_path / "kat")
    subprocess.check_call([hipcc, "-O1", src, "-o", exe], cwd=str(tmp_path))
    rng = np.random.default_rng(5)
    rows = rng.integers(0, 2 ** 32, size=(200, 6), dtype=np.uint64)
    inp = "\n".join(" ".join(str(int(x)) for x in r) for r in rows) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for r, line in zip(rows, out):
        want = [int(x) for x in line.split()]
        got = orc.philox(r[:4], r[4:6])
        assert list(got) == want


def test_u01_open_interval_and_uniformity(orc):
    # u01(x) = (2*(x>>9)+1) * 2^-24: strictly inside (0,1) at both extremes, exact in fp32
    assert orc.u01(0) == np.float32(2.0 ** -24)
    assert orc.u01(0xffffffff) == np.float32(1 - 2.0 ** -24) and orc.u01(0xffffffff) < 1.0
    assert orc.u01(0x80000000) == np.float32(0.5 + 2.0 ** -24)
    words = np.concatenate([orc.philox([i, 4, 0, 0], [1, 2]) for i in range(2000)])
    u = np.array([orc.u01(int(w)) for w in words])
    assert abs(u.mean() - 0.5) < 4 * np.sqrt(1 / 12 / u.size)
    assert abs((u ** 2).mean() - 1 / 3) < 4 * np.sqrt(4 / 45 / u.size)
