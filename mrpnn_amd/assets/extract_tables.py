"""Transcribe the numeric tables the hot path needs from the reference's sources.

Tables (data only, GPL-3.0 like their source):
* the sphere-stencil direction tables Spn0..Spn4 and SpVn5..SpVn7
  (core/vector.cu:82-177) used by SphereRandom3 (core/vector.cu:265-302);
* the prime bases and the first digit-permutation entries of the scrambled Halton
  sequence (core/volume.cu:1189-1199) that Fill_Hg (core/volume.cu:1322-1400) reads
  for dimensions 0, 1, 2 and 4 (bases 2, 3, 5, 11; permutation offsets 0..27).

Writes ``mrpnn_amd/csrc/ref_tables.h``.  ``tests/test_tables.py`` re-parses the
reference and checks the header against it when the reference is present.

Usage:  python mrpnn_amd/assets/extract_tables.py [/path/to/reference]
"""
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "csrc", "ref_tables.h")
NAMES = ["Spn0", "Spn1", "Spn2", "Spn3", "Spn4", "SpVn5", "SpVn6", "SpVn7"]


def parse_stencil(vector_cu):
    tables = {}
    for name in NAMES:
        m = re.search(name + r"\[(\d+)\]\s*=\s*\{(.*?)\};", vector_cu, re.S)
        n = int(m.group(1))
        trip = re.findall(r"float3\{([^}]*)\}", m.group(2))
        vals = [tuple(float(x) for x in t.split(",")) for t in trip]
        assert len(vals) == n, name
        tables[name] = vals
    return tables


def parse_halton(volume_cu):
    primes = [int(x) for x in re.search(r"Primes\[PrimeTableSize\]\s*=\s*\{([^}]*)\}",
                                          volume_cu).group(1).replace("\n", " ").split(",") if x.strip()]
    sums = [int(x) for x in re.search(r"PrimeSums\[PrimeTableSize\]\s*=\s*\{([^}]*)\}",
                                        volume_cu).group(1).replace("\n", " ").split(",") if x.strip()]
    perm = [int(x) for x in re.search(r"RadicalInversePermutations_\[\]\s*=\s*\{([^}]*)\}",
                                        volume_cu, re.S).group(1).replace("\n", " ").split(",") if x.strip()]
    return primes, sums, perm


def main():
    vec = open(os.path.join(REF, "core", "vector.cu")).read()
    vol = open(os.path.join(REF, "core", "volume.cu")).read()
    st = parse_stencil(vec)
    primes, sums, perm = parse_halton(vol)
    lines = ["// Generated by mrpnn_amd/assets/extract_tables.py -- data tables of sim00000n/MRPNN",
             "// (GPL-3.0).  Do not edit by hand.",
             "#pragma once", "", "namespace mrp {", "",
             "// Stencil directions in the (XMain, YMain, ZMain) frame, 160 entries:",
             "// Spn0[8] Spn1[8] Spn2[16] Spn3[16] Spn4[16] SpVn5[32] SpVn6[32] SpVn7[32]",
             "// (core/vector.cu:82-177).",
             "static const float kStencilDirs[160][3] = {"]
    for name in NAMES:
        lines.append("    // " + name)
        for v in st[name]:
            lines.append("    {%sf, %sf, %sf}," % tuple(repr(float(x)) for x in v))
    lines.append("};")
    lines.append("")
    lines.append("// Halton bases and digit permutations for dimensions 0..4 (core/volume.cu:1189-1199).")
    need = sums[4] + primes[4]
    lines.append("// Initialiser macros so device code can declare __constant__ copies.")
    lines.append("#define MRP_HALTON_PRIMES_INIT {%s}" % ", ".join(str(p) for p in primes[:5]))
    lines.append("#define MRP_HALTON_PRIME_SUMS_INIT {%s}" % ", ".join(str(s) for s in sums[:5]))
    lines.append("#define MRP_HALTON_PERM_INIT {%s}" % ", ".join(str(p) for p in perm[:need]))
    lines.append("static const unsigned kHaltonPrimes[5] = MRP_HALTON_PRIMES_INIT;")
    lines.append("static const unsigned kHaltonPrimeSums[5] = MRP_HALTON_PRIME_SUMS_INIT;")
    lines.append("static const unsigned kHaltonPerm[%d] = MRP_HALTON_PERM_INIT;" % need)
    lines += ["", "}  // namespace mrp", ""]
    open(OUT, "w").write("\n".join(lines))
    print("wrote", os.path.normpath(OUT), "perm entries", need)


if __name__ == "__main__":
    main()
