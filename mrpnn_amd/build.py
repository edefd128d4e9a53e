"""Build recipe of libmrpnn_hip.so (gfx950) and of the C++ host API objects.

Every translation unit under mrpnn_amd/csrc is compiled with hipcc for gfx950 in-tree
(objects in mrpnn_amd/_build/, library next to this file), in parallel.

    python -m mrpnn_amd.build            # or build() from __graft_entry__
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmrpnn_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MRP_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: float expressions round as written (FMAs only where written);
# packed fp32 ops are disabled for all device code.
# -amdgpu-mfma-vgpr-form: MFMA results in VGPRs, so the network kernel fits three waves per
# SIMD (accumulation registers would otherwise be allocated on top of the VGPR budget).
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]

# Per-file additions.  mlp_mfma.hip: no input of its network kernels is NaN, and without NaN
# semantics a ReLU / max is one v_max_f32 (IEEE mode would quiet every MFMA result first).
# Scheduling: the AMDGPU register-pressure trackers and the memory-clause strategy keep more of
# the network kernels' gather loads together and spill less (1080p network kernel 0.996 ->
# 0.983 ms); the march also takes the clause strategy.  They change instruction order and
# register assignment only, so results are bit-identical.
_CLAUSE = ["-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]
FILE_FLAGS = {"mlp_mfma.hip": ["-fno-honor-nans", "-mllvm", "-amdgpu-use-amdgpu-trackers=1"] + _CLAUSE,
              "fused.hip": _CLAUSE}

SOURCES = ["capi.hip", "volume_kernels.hip", "predict_plain.hip", "fused.hip", "fused_host.hip", "mlp_mfma.hip",
           "pathtrace.hip", "tile_gather.hip", "interactive.hip", "rpnn.hip", "ingest.cpp", "serve.cpp", "cxx_api.cpp"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return out, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
