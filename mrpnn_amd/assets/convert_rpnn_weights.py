"""Convert the RPNN comparison network's weights from the reference's C header into a binary
asset, as convert_weights.py does for the MRPNN network.

Source: core/NNWeight_RPNN.cuh of sim00000n/MRPNN (GPL-3.0): 54 fp32 arrays (1,296,401
parameters) of ``RadiancePredict_RPNN`` (core/radiancePredict.cu:567-722): per level l = 0..9
``V<l>W`` [200][226] (no bias), for l = 1..9 ``W<l>W`` [200][200] + ``W<l>B``, for l = 0..9
``U<l>W`` [200][200] + ``U<l>B``, then ``F0``, ``F1`` ([200][200] + bias) and ``F2``
([1][200] + bias); row-major ``W[out][in]`` (``Weight[j + i * From]``).

The decimal text is parsed (nothing from the header is executed or compiled) into
``rpnn_weights.bin`` (float32, little endian, header order) and ``rpnn_weights.json`` (name,
offset, count of every array and the SHA-256 of the .bin).  GPL-3.0 data derived from the
reference.

Usage:  python mrpnn_amd/assets/convert_rpnn_weights.py [path/to/NNWeight_RPNN.cuh]
"""
import hashlib
import json
import os
import struct
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from convert_weights import parse  # noqa: E402

DEFAULT_SRC = "/root/reference/core/NNWeight_RPNN.cuh"


def expected_layout():
    names = [("V%dW" % l, 200 * 226) for l in range(10)]
    for l in range(1, 10):
        names += [("W%dW" % l, 40000), ("W%dB" % l, 200)]
    for l in range(10):
        names += [("U%dW" % l, 40000), ("U%dB" % l, 200)]
    names += [("F0W", 40000), ("F0B", 200), ("F1W", 40000), ("F1B", 200), ("F2W", 200), ("F2B", 1)]
    return names


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    with open(src, "r") as f:
        arrays = parse(f.read())
    if [(n, len(v)) for n, v in arrays] != expected_layout():
        raise ValueError("unexpected array order or sizes in " + src)
    here = os.path.dirname(os.path.abspath(__file__))
    blob = bytearray()
    manifest = {"source": "sim00000n/MRPNN core/NNWeight_RPNN.cuh", "license": "GPL-3.0",
                "dtype": "float32-le", "arrays": []}
    off = 0
    for name, vals in arrays:
        blob += struct.pack("<%df" % len(vals), *vals)
        manifest["arrays"].append({"name": name, "offset": off, "count": len(vals)})
        off += len(vals)
    manifest["total"] = off
    manifest["sha256"] = hashlib.sha256(bytes(blob)).hexdigest()
    with open(os.path.join(here, "rpnn_weights.bin"), "wb") as f:
        f.write(bytes(blob))
    with open(os.path.join(here, "rpnn_weights.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(f"{len(arrays)} arrays, {off} floats, sha256 {manifest['sha256']}")


if __name__ == "__main__":
    main()
