"""Calibrates the FP8 activation exponents that ship with the library (csrc/fp8_scales.h).

Max-abs calibration: the BF16 network runs over the features of every active pixel of a set of
seeded fBm views (mrp_fp8_calibrate) and records, per MFMA layer, the largest |input|; a layer's
exponent is the largest e with max * 2^e <= 448.  The views cover the benchmark workloads:
512^3 clouds (6 octaves, seeds 1 and 2) with alpha 64 and 16 suns, and 128^3 / 256^3 / 384^3
clouds with alpha 16, 64 and 128; albedos 0.99 and 0.999 and one view with three distinct
colour channels (whose tails then run per channel).  Needs a GPU.

    python -m mrpnn_amd.assets.calibrate_fp8 --out mrpnn_amd/csrc/fp8_scales.h
"""
import argparse
import sys
import time

import numpy as np

G = 0.857


def sun_dirs(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        el = rng.uniform(np.radians(10), np.radians(80))
        az = rng.uniform(0, 2 * np.pi)
        out.append((np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)))
    return out


def workload():
    """(resolution, octaves, seed, alpha, albedo3, sun, azimuth, elevation) of every calibration view."""
    views = []
    for seed in (1, 2):
        for i, L in enumerate(sun_dirs(16, 100 + seed)):
            views.append((512, 6, seed, 64.0, (0.99, 0.99, 0.99), L, 0.6 + 0.4 * i, 0.25))
    rng = np.random.default_rng(7)
    for res in (128, 256, 384):
        for alpha in (16.0, 64.0, 128.0):
            L = sun_dirs(1, int(rng.integers(1 << 30)))[0]
            views.append((res, 6, 3, alpha, (0.999, 0.999, 0.999), L, float(rng.uniform(0, 6.28)), 0.3))
    views.append((256, 6, 4, 64.0, (0.9, 0.95, 0.99), sun_dirs(1, 5)[0], 1.0, 0.2))
    return views


def calibrate(width=512, height=512, log=print):
    import torch

    import mrpnn_amd as M
    ctx = M.Context()
    vols = {}
    first = True
    t0 = time.perf_counter()
    rows = 0
    for (res, octv, seed, alpha, alb, L, az, el) in workload():
        key = (res, octv, seed)
        if key not in vols:
            for v in vols.values():
                v.close()
            vols.clear()
            vol = M.Volume(ctx, res)
            vol.fill_fbm(octv, 4.0, seed)
            vol.update()
            vols[key] = vol
        vol = vols[key]
        v = M.orbit_view(width, height, dist=2.0, azimuth=az, elevation=el, fov_deg=40.0, light=L, alpha=alpha, g=G,
                         albedo=0.99, seed=seed, jitter=0.5)
        v.scatter_rate[:] = list(alb)
        ctx.prepare_views([vol], [v])  # LUT for g
        vol.update_tr(np.asarray(v.light_dir[:], np.float32), v.alpha)
        feats, active, _, _ = vol.gather_features(v)
        f = feats[active > 0].contiguous()
        sr = [a / 1.001 for a in alb]
        maxima = ctx.fp8_calibrate(f, sr, G, reset=first)
        first = False
        rows += f.shape[0]
        del feats, f
        torch.cuda.synchronize()
        log("  %d^3 seed %d alpha %g: %d rows so far (%.1f s)" % (res, seed, alpha, rows, time.perf_counter() - t0))
    exps = ctx.fp8_exponents()
    for v in vols.values():
        v.close()
    ctx.close()
    log("calibrated on %d feature rows in %.1f s" % (rows, time.perf_counter() - t0))
    return maxima, exps, rows


HEADER = """// FP8 activation exponents of the MFMA network's 60 layers (mlp_mfma.h, MfmaLayerId order):
// the inputs of layer id are stored as e4m3(clamp(x, +-448 * 2^-e) * 2^e), e = kFp8ActExp[id],
// the largest e with (max |input|) * 2^e <= 448.  Calibrated by mrpnn_amd/assets/calibrate_fp8.py
// (BF16 network over %d active-pixel feature rows of seeded fBm views: 512^3 clouds, alpha 64,
// 32 suns; 128^3-384^3 clouds, alpha 16-128; albedo 0.99 / 0.999 / distinct channels).
#pragma once

namespace mrp {
// per layer: max |input| seen in calibration, then the exponent
constexpr int kFp8ActExp[60] = {
%s
};
}  // namespace mrp
"""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="header to write (default: print)")
    ap.add_argument("--res", type=int, default=512, help="view width and height")
    a = ap.parse_args()
    maxima, exps, rows = calibrate(a.res, a.res, log=lambda s: print(s, file=sys.stderr, flush=True))
    lines = ["    %4d,  // layer %2d: max |input| %.6g" % (e, i, m) for i, (e, m) in enumerate(zip(exps, maxima))]
    text = HEADER % (rows, "\n".join(lines))
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)
    else:
        print(text)


if __name__ == "__main__":
    main()
