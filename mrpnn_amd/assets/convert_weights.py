"""Convert the MRPNN network weights from the reference's C header into a binary asset.

Source: core/NNWeight.cuh of sim00000n/MRPNN (GPL-3.0).  That header holds 146 fp32
arrays (49,673 parameters) of the multi-feature fusion MLP evaluated by
``RadiancePredictInternal`` (core/radiancePredict.cu:301-422).  Weight matrices are
row-major ``W[out][in]`` (``Weight[j + i * From]``, core/radiancePredict.cu:51).

The script parses the decimal text (no code from the header is executed or compiled)
and writes, next to itself:

* ``mrpnn_weights.bin``  -- the 49,673 float32 values, little endian, concatenated in
  header order;
* ``mrpnn_weights.json`` -- manifest: name, element offset and count of every array,
  plus the SHA-256 of the ``.bin`` file.

The output is GPL-3.0 data derived from the reference and carries that licence.

Usage:  python mrpnn_amd/assets/convert_weights.py [path/to/NNWeight.cuh]
"""
import hashlib
import json
import os
import re
import struct
import sys

DEFAULT_SRC = "/root/reference/core/NNWeight.cuh"
DECL = re.compile(
    r"__device__\s+(?:__constant__\s+)?(?:const\s+)?float\s+(\w+)\[(\d+)\]\s*=\s*\{([^}]*)\}")


def parse(text):
    arrays = []
    for name, count, body in DECL.findall(text):
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        if len(vals) != int(count):
            raise ValueError(f"{name}: declared {count} values, found {len(vals)}")
        arrays.append((name, vals))
    return arrays


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    with open(src, "r") as f:
        arrays = parse(f.read())
    here = os.path.dirname(os.path.abspath(__file__))
    blob = bytearray()
    manifest = {"source": "sim00000n/MRPNN core/NNWeight.cuh", "license": "GPL-3.0",
                "dtype": "float32-le", "arrays": []}
    off = 0
    for name, vals in arrays:
        blob += struct.pack("<%df" % len(vals), *vals)
        manifest["arrays"].append({"name": name, "offset": off, "count": len(vals)})
        off += len(vals)
    manifest["total"] = off
    manifest["sha256"] = hashlib.sha256(bytes(blob)).hexdigest()
    with open(os.path.join(here, "mrpnn_weights.bin"), "wb") as f:
        f.write(bytes(blob))
    with open(os.path.join(here, "mrpnn_weights.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    write_layout(manifest, bytes(blob), os.path.join(here, "..", "csrc", "net_layout.h"))
    print(f"{len(arrays)} arrays, {off} floats, sha256 {manifest['sha256']}")


def fnv1a64(data):
    h = 0xcbf29ce484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def write_layout(manifest, blob, path):
    """Element offsets of every array in the asset, for the HIP library."""
    lines = ["// Generated by mrpnn_amd/assets/convert_weights.py: offsets (in floats) of the",
             "// MRPNN weight arrays inside mrpnn_weights.bin (order of core/NNWeight.cuh).",
             "#pragma once", "", "namespace mrp {", "namespace net {"]
    for a in manifest["arrays"]:
        lines.append("constexpr int %s = %d;  // %d" % (a["name"], a["offset"], a["count"]))
    lines.append("constexpr int TOTAL = %d;" % manifest["total"])
    lines.append("constexpr unsigned long long FNV1A64 = 0x%016xull;  // of the .bin bytes" % fnv1a64(blob))
    lines += ["}  // namespace net", "}  // namespace mrp", ""]
    open(path, "w").write("\n".join(lines))


if __name__ == "__main__":
    main()
