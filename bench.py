"""Benchmark of the MRPNN radiance predictor on one MI355X.

Default workload (the headline metric): 1920x1080 frames over a 512^3 seeded fBm cloud
(6 octaves, base frequency 4, density scale alpha = 64, g = 0.857, albedo 0.99), camera at
distance 2 with a 40 degree field of view, the sun cycling through 16 seeded directions on
the upper hemisphere.  One step = one frame: TR-mip rebake for the new sun (as the
reference's Render does on a light change) + march + gather + MLP + composition, inputs
resident in HBM.  `value` is the BF16 fused path in Mpix/s (frame pixels / wall time);
the plain per-frame FP32 kernel (the drop-in baseline) and the FP8 path are timed on the
same frames.  p50/p99 are per-frame latencies of the BF16 path (HIP events on the
library's stream).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD every 4 cycles at the
# 2.4 GHz peak engine clock (wave64 instructions per second / 1e9)
VALU_PEAK_GINSTR = 256 * 4 * 2.4 / 4
FLOP_PER_PIXEL = 97072  # 2 x 48,536 MAC per pixel, one distinct colour channel
GATHER_BYTES_PER_PIXEL = 18412  # logical fp32 texel bytes of the 192-point gather (+ LUT pairs, output)


def sun_dirs(n, seed=11):
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        z = rng.uniform(0.15, 0.95)  # elevation sine (upper hemisphere, y up)
        phi = rng.uniform(0, 2 * math.pi)
        r = math.sqrt(1 - z * z)
        d = np.array([r * math.cos(phi), z, r * math.sin(phi)], np.float32)
        out.append(d / np.linalg.norm(d))
    return out


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--tile-w", type=int, default=8, help="march tile width of the batched path (4..64)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy 1 GiB device-to-device (scale reference for HBM counters)")
    ap.add_argument("--config", default="headline", choices=["headline", "c1", "c2", "c3", "c4", "c5", "c5sweep", "tile", "interactive", "ingest",
                             "rpnn"],
                    help="headline = the metric line; c1..c5, tile, interactive, ingest, rpnn = the other workloads (one JSON line each)")
    ap.add_argument("--pt-res", type=int, default=1024, help="c3: image side")
    ap.add_argument("--pt-spp", type=int, default=1024, help="c3: samples per pixel")
    ap.add_argument("--pt-bounces", type=int, default=512, help="c3: maximum bounces")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="c5: arrival window")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="c5: offered requests per second")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mrpnn_amd as M

    torch.cuda.set_device(0)
    dev = torch.device("cuda")
    if args.copy_probe:
        src = torch.ones(256 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 1 GiB read + 1 GiB written")
    if args.config != "headline":
        import bench_configs
        print(json.dumps(getattr(bench_configs, args.config)(args)))
        return
    ctx = M.Context()
    ctx.set_timing(True)
    # one cached TR set, as the reference keeps (volume.cu:1592): the sun moves every frame,
    # so every frame rebakes its TR mips (a cache of 16 or more would serve the 16 cycling suns
    # from memory).  The bake runs on the library's second stream beside the march; with two
    # sets it could also overlap the previous frame's network, which measured slower (both are
    # VALU-bound and share the CUs)
    ctx.set_option(M.OPT_TR_SLOTS, 1)
    ctx.set_option(M.OPT_TILE_W, args.tile_w)
    W, H = args.width, args.height
    vol = M.Volume(ctx, args.res)
    t0 = time.time()
    vol.fill_fbm(6, 4.0, 1)
    torch.cuda.synchronize()
    log("fbm %d^3 generated in %.2fs" % (args.res, time.time() - t0))
    torch.cuda.synchronize()
    t0 = time.time()
    vol.update()
    torch.cuda.synchronize()
    pyramid_ms = (time.time() - t0) * 1e3
    log("pyramid build %.2f ms" % pyramid_ms)

    suns = sun_dirs(16)

    def view(i):
        return M.orbit_view(W, H, dist=2.0, azimuth=0.6, elevation=0.25, fov_deg=40.0, light=suns[i % 16],
                            alpha=64.0, g=0.857, albedo=0.99, seed=1, frame=0, view_id=0, jitter=0.5)

    out = torch.zeros((H * W, 4), device=dev)

    def frame(i, prec):
        v = view(i)
        if prec == "plain":  # rebuilds LUT / TR mips itself when the sun changed
            vol.render_plain_device(v, out)
        else:
            ctx.prepare_views([vol], [v])
            ctx.render_batch([vol], [v], [out], M.PRECISIONS[prec])

    results = {}
    kstats = {}
    for prec, steps in (("bf16", args.steps), ("fp8", args.steps), ("plain", args.plain_steps)):
        for i in range(args.warmup if prec != "plain" else 1):
            frame(i, prec)
        torch.cuda.synchronize()
        ctx.set_timing(True)  # resets the device-time sums; timing never blocks a frame
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record()
        for i in range(steps):
            frame(args.warmup + i, prec)
            ev[i + 1].record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
        s = ctx.stats()
        assert s.timed_renders == steps, (s.timed_renders, steps)
        results[prec] = dict(wall=wall, steps=steps, lat=lat, mpix=W * H * steps / wall / 1e6)
        kstats[prec] = dict(march_ms=s.ms_march_sum / steps, predict_ms=s.ms_predict_sum / steps,
                            active=s.active_pixels, taps=s.march_taps)
        log("%-5s %7.1f Mpix/s  %.2f ms/frame  p50 %.2f p99 %.2f ms  (march %.3f ms, gather+mlp %.3f ms)" %
            (prec, results[prec]["mpix"], 1e3 * wall / steps, np.percentile(lat, 50), np.percentile(lat, 99),
             kstats[prec]["march_ms"], kstats[prec]["predict_ms"]))

    # accuracy of the BF16 / FP8 frames against the FP32 plain kernel (the reference's
    # arithmetic) on the same view: mean relative radiance error over pixels brighter than
    # 1e-3 of the frame maximum (per channel)
    def image(prec):
        frame(0, prec)
        torch.cuda.synchronize()
        return out.clone()
    ref = image("plain")
    act = ref[:, 3] > 0
    n_active = int(act.sum().item())
    rgb = ref[:, :3]
    bright = rgb > 1e-3 * rgb.amax(dim=0, keepdim=True)
    errs = {}
    for prec in ("bf16", "fp8"):
        img = image(prec)[:, :3]
        rel = ((img - rgb).abs() / rgb.clamp_min(1e-30))[bright]
        errs[prec] = dict(mean=float(rel.mean().item()), p99=float(torch.quantile(rel.float()[:1 << 24], 0.99).item()))
    log("active pixels %d (%.1f%%); bf16 mean rel err %.5f (p99 %.4f), fp8 %.4f (p99 %.3f)" %
        (n_active, 100.0 * n_active / (W * H), errs["bf16"]["mean"], errs["bf16"]["p99"], errs["fp8"]["mean"],
         errs["fp8"]["p99"]))

    # roofline of the dominant kernel (fused gather + BF16 MFMA network): its average device
    # time over the timed frames, from the library's HIP events around that launch on its stream
    ks = kstats["bf16"]
    pred_ms = ks["predict_ms"]
    march_ms = ks["march_ms"]
    taps = float(ks["taps"])
    flops = FLOP_PER_PIXEL * n_active
    achieved_tf = flops / (pred_ms * 1e-3) / 1e12
    gather_gbs = GATHER_BYTES_PER_PIXEL * n_active / (pred_ms * 1e-3) / 1e9
    march_gbs = 32.0 * taps / (march_ms * 1e-3) / 1e9
    traffic = None
    valu = None
    tpath = os.path.join(ROOT, "profiles", "traffic_mlp_mfma_active_bf16_1080p_512.json")
    if os.path.exists(tpath):
        prof = json.load(open(tpath))
        traffic = prof.get("hbm_bytes_per_launch")
        per_tile = prof.get("valu_instr_per_16px_tile")
        if per_tile:
            # the network kernel issues VALU instructions, not MFMAs, at its bound: rocprof's VALU
            # instruction count per 16-pixel tile over this run's kernel time, against the issue peak
            g = per_tile * (n_active / 16.0) / (pred_ms * 1e-3) / 1e9
            valu = {"bound": "valu", "achieved": round(g, 1), "peak": VALU_PEAK_GINSTR, "unit": "Gwave-instr/s",
                    "frac": round(g / VALU_PEAK_GINSTR, 4), "valu_instr_per_tile": per_tile}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(vol, ctx, view(args.warmup), args.cpu_seconds, W, H)

    bf = results["bf16"]
    line = {
        "metric": "Mpix/s predictor @1080p, 512^3 cloud (bf16 vs fp32) + p50/p99 frame latency",
        "value": round(bf["mpix"], 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * bf["wall"] / bf["steps"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic: seeded 6-octave fBm cloud, 16 seeded sun directions",
        "config": {"workload": "1920x1080 single view over a 512^3 fBm cloud, sun swept over 16 directions, "
                               "TR rebake per frame, BF16 MFMA network", "width": W, "height": H,
                   "volume": args.res, "alpha": 64.0, "g": 0.857, "albedo": 0.99, "fov_deg": 40.0},
        "p50_ms": round(float(np.percentile(bf["lat"], 50)), 3),
        "p99_ms": round(float(np.percentile(bf["lat"], 99)), 3),
        "fp32_plain_mpix_s": round(results["plain"]["mpix"], 2),
        "fp8_mpix_s": round(results["fp8"]["mpix"], 2),
        "speedup_bf16_vs_fp32_plain": round(bf["mpix"] / results["plain"]["mpix"], 2),
        "bf16_mean_rel_err_vs_fp32": round(errs["bf16"]["mean"], 5),
        "bf16_p99_rel_err_vs_fp32": round(errs["bf16"]["p99"], 5),
        "fp8_mean_rel_err_vs_fp32": round(errs["fp8"]["mean"], 5),
        "active_pixel_fraction": round(n_active / (W * H), 4),
        "pyramid_build_ms": round(pyramid_ms, 2),
        "kernel_ms": {"march_compact": round(march_ms, 3), "mlp_mfma_active_bf16": round(pred_ms, 3),
                      "mlp_mfma_active_fp8": round(kstats["fp8"]["predict_ms"], 3),
                      "predict_plain_fp32": round(kstats["plain"]["predict_ms"], 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved_tf, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tf / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": traffic},
        "roofline_valu": valu,
        "roofline_gather": {"bound": "hbm", "achieved": round(gather_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gather_gbs / HBM_PEAK_GBS, 4),
                            "note": "logical fp32 texel bytes of the gather; > 1 means on-chip reuse"},
        "roofline_march": {"bound": "hbm", "achieved": round(march_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(march_gbs / HBM_PEAK_GBS, 4), "taps_per_frame": int(taps)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def cpu_baseline(vol, ctx, v, seconds, W, H):
    """The oracle's restatement of NNPredict<MRPNN> on the host cores (OpenMP over pixels),
    on a seeded uniform sample of the same frame's pixels, over the same pyramids."""
    import numpy as np
    import torch

    import oracle as O
    from bench_configs import cpu_info, cpu_threads
    threads = cpu_threads()
    O.set_threads(threads)
    ctx.prepare_views([vol], [v])
    vol.update_tr(np.asarray(v.light_dir[:], np.float32), v.alpha)  # the volume's own TR mips for this sun
    torch.cuda.synchronize()
    sc = O.Scene.from_arrays(vol.get_data(), vol.N, 1.0, vol.download_mips(), vol.download_tr(),
                             ctx.download_hg_lut())
    ov = O.View.from_buffer_copy(bytes(v))
    rng = np.random.default_rng(0)
    n = 256
    t0 = time.perf_counter()
    sc.render_predict_list(ov, rng.integers(0, W * H, n))
    dt = time.perf_counter() - t0
    n = int(max(256, min(W * H, n * seconds / max(dt, 1e-3))))
    pix = rng.integers(0, W * H, n)
    t0 = time.perf_counter()
    sc.render_predict_list(ov, pix)
    dt = time.perf_counter() - t0
    log("cpu baseline: %d pixels in %.2fs on %d threads" % (n, dt, threads))
    return dict({"value": round(n / dt / 1e6, 5), "unit": "Mpix/s", "cores": threads, "kind": "port",
                 "sample": "%d pixels drawn uniformly from the benchmark 1080p frame, same pyramids (%.1f s)" % (n, dt)},
                **cpu_info())


if __name__ == "__main__":
    main()
