"""The workloads beside the headline metric, one JSON line each (`python bench.py --config cN`).

c1  plain FP32 and BF16 predictor at 256x256 over a 128^3 cloud, the oracle's CPU predictor
    on the host cores, and the delta-tracking judge (256 spp, <= 64 bounces);
c2  precision sweep at 1920x1080 over 512^3 clouds (seeds 1..4, 16 suns): FP32 / BF16 / FP8
    against the FP64 network on identical features (sampled pixels) and against FP32;
c3  delta-tracking judge at 1024^2, 1024 spp, <= 512 bounces, albedo 0.999: samples/s and
    the fraction of predictor pixels inside the 3-sigma band;
c4  64 views over 8 resident volumes (two each of 128^3, 256^3, 384^3, 512^3), resolutions
    from {256^2, 512^2, 1280x720, 1920x1080}, one graph-replayed BF16 batch per step with
    the autotuned launch shape;
c5  serving loop: Poisson arrivals over 32 resident volumes, BF16 / FP8 batched against
    the plain per-frame FP32 launch;
tile  tile feature gather (the training-input path, MRPNNInputData.fetch_input) at 1920x1080
    over a 512^3 cloud, with the oracle's CPU gather of the same frame as baseline and check;
interactive  the GUI's progressive renderer at 1024x1024 over a 512^3 cloud (MRPNN BF16 / FP32,
    path tracing with and without the histogram filter);
ingest  loaders (.vox / .txt / .bin), the RGBE reader and the cost of an environment map;
rpnn  the RPNN comparison predictor at 1920x1080 over a 512^3 cloud (FP32 and BF16 MFMA paths),
    against the MRPNN BF16 frame of the same view and the oracle's RPNN on the host cores.

All volumes are the seeded fBm clouds of the library (mrp_volume_fill_fbm); the oracle is the
CPU restatement under oracle/ and is used here only as the timed CPU baseline and as the FP64
reference network.
"""
import math
import os
import time

import numpy as np

G = 0.857


def cpu_threads():
    """Threads of the CPU baselines: the host CPU share this GPU is given (OMP_NUM_THREADS, which
    the GPU box sets to 16 per GPU; MRP_CPU_THREADS overrides it), never more than the CPUs this
    process may run on; all of those CPUs when neither variable is set."""
    aff = len(os.sched_getaffinity(0))
    env = int(os.environ.get("MRP_CPU_THREADS", "0") or 0) or int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(env, aff) if env > 0 else aff


def cpu_info():
    """Host CPU description for the baseline records: model name, logical CPUs, affinity."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "thread_rule": "OMP_NUM_THREADS (the host CPU share of this GPU), capped by the affinity mask"}


def _log(*a):
    import sys
    print("[bench]", *a, file=sys.stderr, flush=True)


def _sun(elev_deg, az_deg):
    e, a = math.radians(elev_deg), math.radians(az_deg)
    return (math.cos(e) * math.cos(a), math.sin(e), math.cos(e) * math.sin(a))


def _band(pred, mean, m2, n):
    """Fraction of pixels (channel 0) with |pred - mean| <= 3 * standard error."""
    se = np.sqrt(np.maximum(m2[:, 0], 0.0) / (n * (n - 1)))
    return float(np.mean(np.abs(pred[:, 0] - mean[:, 0]) <= 3.0 * se))


def _judge(pred, mean, m2, n, W, H, taps=None):
    """The per-pixel 3-sigma judge of the predictor against the delta-tracking image, with where
    the failing pixels sit: (a) pixels the predictor leaves as sky (no first-scattering point:
    survival never drops below 0.5 along the ray) while the path tracer sees light; (b) among
    predicted pixels, by radiance decile of the path-traced mean, by distance (pixels) to the
    predictor's silhouette, and by march steps to the hit; (c) the signed relative bias."""
    from scipy import ndimage
    p, mu = np.asarray(pred, np.float64)[:, 0], np.asarray(mean, np.float64)[:, 0]
    se = np.sqrt(np.maximum(np.asarray(m2, np.float64)[:, 0], 0.0) / (n * (n - 1)))
    act = np.asarray(pred)[:, 3] > 0
    ok = np.abs(p - mu) <= 3.0 * se
    lit = mu > 1e-4 * mu.max()
    res = {"pixels": int(p.size), "pass_all": round(float(ok.mean()), 4),
           "sky_in_predictor_but_lit": int((~act & lit & ~ok).sum()),
           "sky_in_predictor_fail_share": round(float((~act & ~ok).sum() / max((~ok).sum(), 1)), 4),
           "pass_predicted": round(float(ok[act].mean()), 4) if act.any() else None}
    a = act & lit
    rel = (p[a] - mu[a]) / mu[a]
    res["predicted_bias"] = {"mean_signed_rel": round(float(rel.mean()), 4),
                             "mean_abs_rel": round(float(np.abs(rel).mean()), 4),
                             "median_se_rel": round(float(np.median(se[a] / mu[a])), 4)}
    dec = np.quantile(mu[a], np.linspace(0, 1, 11))
    idx = np.clip(np.searchsorted(dec, mu[a], side="right") - 1, 0, 9)
    res["fail_by_radiance_decile"] = [round(float(1 - ok[a][idx == k].mean()), 3) for k in range(10)]
    res["signed_rel_by_radiance_decile"] = [round(float(rel[idx == k].mean()), 3) for k in range(10)]
    dist = ndimage.distance_transform_edt(act.reshape(H, W)).ravel()[a]
    bands = [(1, 2), (2, 4), (4, 8), (8, 16), (16, 1e9)]
    res["fail_by_silhouette_distance_px"] = {"%g-%g" % b: round(float(1 - ok[a][(dist >= b[0]) & (dist < b[1])].mean()), 3)
                                            for b in bands if ((dist >= b[0]) & (dist < b[1])).any()}
    if taps is not None:
        tp = np.asarray(taps)[a]
        q = np.quantile(tp, [0, 0.25, 0.5, 0.75, 1.0])
        res["fail_by_march_steps_quartile"] = [round(float(1 - ok[a][(tp >= q[k]) & (tp <= q[k + 1])].mean()), 3)
                                               for k in range(4)]
    return res


def _rel(a, b, floor_frac=1e-3):
    """Mean and p99 relative error of a against b over pixels brighter than floor_frac * max."""
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    m = b > floor_frac * b.max() if b.size and b.max() > 0 else np.zeros_like(b, bool)
    if not m.any():
        return dict(mean=None, p99=None)
    r = np.abs(a[m] - b[m]) / b[m]
    return dict(mean=round(float(r.mean()), 6), p99=round(float(np.quantile(r, 0.99)), 5))


def c1(args):
    import torch

    import mrpnn_amd as M
    import oracle as O
    ctx = M.Context()
    N = 128
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, 1)
    vol.update()
    v = M.orbit_view(256, 256, dist=2.0, azimuth=0.0, elevation=0.0, fov_deg=40.0, light=_sun(30, 0), alpha=64.0,
                     g=G, albedo=0.99, seed=1, jitter=0.5)
    out = torch.zeros((256 * 256, 4), device="cuda")
    res = {}
    for name, fn in (("fp32_plain", lambda: vol.render_plain_device(v, out)),
                     ("bf16", lambda: (ctx.prepare_views([vol], [v]), ctx.render_batch([vol], [v], [out], M.BF16)))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 20
        fn()
        torch.cuda.synchronize()
        res[name] = dict(img=out.cpu().numpy().copy(), mpix=256 * 256 / dt / 1e6)
    # CPU baseline: the oracle's NNPredict<MRPNN> restatement on all host cores, full frame
    threads = cpu_threads()
    O.set_threads(threads)
    vol.update_tr(np.asarray(v.light_dir[:], np.float32), v.alpha)
    torch.cuda.synchronize()
    sc = O.Scene.from_arrays(vol.get_data(), N, 1.0, vol.download_mips(), vol.download_tr(), ctx.download_hg_lut())
    ov = O.View.from_buffer_copy(bytes(v))
    t0 = time.perf_counter()
    cpu = sc.render_predict(ov)["out"]
    cpu_s = time.perf_counter() - t0
    # judge: delta tracking, 256 spp, <= 64 bounces
    t0 = time.perf_counter()
    mean, m2 = vol.pt_render(v, 256, 64, chunk=64)
    torch.cuda.synchronize()
    pt_s = time.perf_counter() - t0
    mean, m2 = mean.cpu().numpy(), m2.cpu().numpy()
    ctx.close()
    return {
        "config": "c1", "workload": "256x256 view of a 128^3 fBm cloud (5 octaves, seed 1), sun 30 deg, alpha 64",
        "fp32_plain_mpix_s": round(res["fp32_plain"]["mpix"], 2), "bf16_mpix_s": round(res["bf16"]["mpix"], 2),
        "cpu_predictor": {"mpix_s": round(256 * 256 / cpu_s / 1e6, 5), "cores": threads, "cpu": cpu_info(), "kind": "port",
                          "seconds": round(cpu_s, 3)},
        "gpu_fp32_vs_cpu_oracle": _rel(res["fp32_plain"]["img"][:, 0], cpu[:, 0]),
        "bf16_vs_fp32": _rel(res["bf16"]["img"][:, 0], res["fp32_plain"]["img"][:, 0]),
        "pt": {"spp": 256, "max_bounce": 64, "seconds": round(pt_s, 3),
               "fp32_in_3sigma": _band(res["fp32_plain"]["img"], mean, m2, 256),
               "bf16_in_3sigma": _band(res["bf16"]["img"], mean, m2, 256),
               "judge_fp32": _judge(res["fp32_plain"]["img"], mean, m2, 256, 256, 256),
               "fp32_vs_pt": _rel(res["fp32_plain"]["img"][:, 0], mean[:, 0])},
    }


def c2(args):
    import torch

    import bench
    import mrpnn_amd as M
    import oracle as O
    threads = cpu_threads()
    O.set_threads(threads)
    ctx = M.Context()
    suns = bench.sun_dirs(16)
    W, H = 1920, 1080
    rng = np.random.default_rng(0)
    acc = {p: {"vs_f64": [], "vs_f32": []} for p in ("fp32", "bf16", "fp8")}
    frames = 0
    t_start = time.perf_counter()
    for seed in (1, 2, 3, 4):
        vol = M.Volume(ctx, 512)
        vol.fill_fbm(6, 4.0, seed)
        vol.update()
        for si, L in enumerate(suns):
            v = M.orbit_view(W, H, dist=2.0, azimuth=0.6, elevation=0.25, fov_deg=40.0, light=L, alpha=64.0, g=G,
                             albedo=0.99, seed=seed, jitter=0.5)
            ctx.prepare_views([vol], [v])  # LUT for g
            vol.update_tr(np.asarray(v.light_dir[:], np.float32), v.alpha)  # the volume's own TR mips
            feats, active, _, _ = vol.gather_features(v)
            act = active > 0
            f = feats[act].contiguous()
            del feats
            sr = np.array([0.99 / 1.001] * 3, np.float32)
            outs = {p: ctx.predict_features(f, sr, G, M.PRECISIONS[p]).cpu().numpy() for p in ("fp32", "bf16", "fp8")}
            idx = rng.choice(f.shape[0], size=min(2048, f.shape[0]), replace=False)
            ref64 = O.predict_f64(f[torch.as_tensor(idx, device="cuda")].cpu().numpy().astype(np.float64),
                                  sr.astype(np.float64), G)
            for p in ("fp32", "bf16", "fp8"):
                acc[p]["vs_f64"].append((outs[p][idx, 0], ref64[:, 0]))
                acc[p]["vs_f32"].append((outs[p][:, 0], outs["fp32"][:, 0]))
            frames += 1
            del f
        _log("c2: seed %d done (%.0f s)" % (seed, time.perf_counter() - t_start))
        vol.close()
    ctx.close()
    res = {"config": "c2", "workload": "1920x1080 views of 512^3 fBm clouds (6 octaves, seeds 1..4), 16 suns",
           "frames": frames, "f64_sample_per_frame": 2048}
    for p in ("fp32", "bf16", "fp8"):
        a64 = np.concatenate([x for x, _ in acc[p]["vs_f64"]]), np.concatenate([y for _, y in acc[p]["vs_f64"]])
        a32 = np.concatenate([x for x, _ in acc[p]["vs_f32"]]), np.concatenate([y for _, y in acc[p]["vs_f32"]])
        res[p] = {"vs_fp64_network": _rel(*a64), "vs_fp32_predictor": _rel(*a32) if p != "fp32" else None}
    return res


def c3(args):
    import torch

    import mrpnn_amd as M
    ctx = M.Context()
    vol = M.Volume(ctx, 512)
    vol.fill_fbm(6, 4.0, 1)
    vol.update()
    R, spp, nb = args.pt_res, args.pt_spp, args.pt_bounces
    v = M.orbit_view(R, R, dist=2.0, azimuth=0.6, elevation=0.25, fov_deg=40.0, light=_sun(30, 0), alpha=64.0,
                     g=G, albedo=0.999, seed=1, jitter=0.5)
    out = torch.zeros((R * R, 4), device="cuda")
    preds = {}
    for p in ("fp32", "bf16"):
        ctx.prepare_views([vol], [v])
        ctx.render_batch([vol], [v], [out], M.PRECISIONS[p])
        torch.cuda.synchronize()
        preds[p] = out.cpu().numpy().copy()
    n_img = torch.zeros((R * R, 3), dtype=torch.float64, device="cuda")
    m2 = torch.zeros_like(n_img)
    done = 0
    t0 = time.perf_counter()
    chunk = 16
    while done < spp:
        k = min(chunk, spp - done)
        vol.pt_render(v, k, nb, chunk=k, mean=n_img, m2=m2, spp0=done)
        torch.cuda.synchronize()
        done += k
        _log("c3: %d / %d spp (%.0f s)" % (done, spp, time.perf_counter() - t0))
    dt = time.perf_counter() - t0
    mean, m2 = n_img.cpu().numpy(), m2.cpu().numpy()
    ctx.close()
    return {"config": "c3", "workload": "%dx%d view of a 512^3 fBm cloud, albedo 0.999, alpha 64" % (R, R),
            "spp": spp, "max_bounce": nb, "seconds": round(dt, 2),
            "samples_per_s": round(R * R * spp / dt, 1),
            "fp32_in_3sigma": _band(preds["fp32"], mean, m2, spp),
            "bf16_in_3sigma": _band(preds["bf16"], mean, m2, spp),
            "judge_fp32": _judge(preds["fp32"], mean, m2, spp, R, R),
            "judge_bf16": _judge(preds["bf16"], mean, m2, spp, R, R),
            "fp32_vs_pt": _rel(preds["fp32"][:, 0], mean[:, 0]),
            "bf16_vs_pt": _rel(preds["bf16"][:, 0], mean[:, 0])}


def c4(args):
    import torch

    import mrpnn_amd as M
    ctx = M.Context()
    ctx.set_timing(True)
    rng = np.random.default_rng(4)
    vols = []
    for k, N in enumerate((128, 128, 256, 256, 384, 384, 512, 512)):
        v = M.Volume(ctx, N)
        v.fill_fbm(6, 4.0, 11 + k)
        v.update()
        vols.append(v)
    sizes = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    views, outs = [], []
    for i in range(64):
        w, h = sizes[rng.integers(0, 4)]
        sun = rng.normal(size=3)
        sun[1] = abs(sun[1]) + 0.15
        views.append(M.orbit_view(w, h, dist=2.0, azimuth=float(rng.uniform(0, 2 * np.pi)),
                                  elevation=float(rng.uniform(-0.3, 0.6)), fov_deg=40.0, light=tuple(sun),
                                  alpha=float(rng.uniform(16, 128)), g=G, albedo=0.99, seed=1, view_id=i,
                                  volume_index=int(rng.integers(0, 8))))
        outs.append(torch.zeros((w * h, 4), device="cuda"))
    pixels = sum(v.width * v.height for v in views)
    t0 = time.perf_counter()
    ctx.prepare_views(vols, views)  # 64 TR sets (one per sun)
    torch.cuda.synchronize()
    bake_ms = (time.perf_counter() - t0) * 1e3
    best_w, best_ms = ctx.autotune(vols, views, outs, M.BF16)
    best_lds = ctx.get_option(M.OPT_LDS_WEIGHTS)
    best_tile = ctx.get_option(M.OPT_TILE_W)
    ctx.set_option(M.OPT_GRAPH, 1)
    res = {}
    for p in ("bf16", "fp8"):
        for _ in range(2):
            ctx.render_batch(vols, views, outs, M.PRECISIONS[p])
        torch.cuda.synchronize()
        steps = 10
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record()
        for i in range(steps):
            ctx.render_batch(vols, views, outs, M.PRECISIONS[p])
            ev[i + 1].record()
        torch.cuda.synchronize()
        lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
        res[p] = {"ms_per_batch": round(float(np.mean(lat)), 3), "mpix_s": round(pixels / np.mean(lat) / 1e3, 1),
                  "p99_batch_ms": round(float(np.percentile(lat, 99)), 3)}
    ctx.set_option(M.OPT_GRAPH, 0)
    # plain per-frame FP32 launches of the same 64 views, for scale
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, v in enumerate(views):
        vols[v.volume_index].render_plain_device(v, outs[i])
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    for v in vols:
        v.close()
    ctx.close()
    return {"config": "c4", "workload": "64 views, 8 volumes (2 x 128^3, 256^3, 384^3, 512^3), mixed resolutions",
            "pixels_per_batch": pixels, "tr_bake_ms_64_suns": round(bake_ms, 1),
            "autotune": {"march_tile": "%dx%d" % (best_tile, 64 // best_tile), "waves_per_cu": best_w,
                         "lds_weights": best_lds, "ms": round(best_ms, 3)}, "graph": True,
            "bf16": res["bf16"], "fp8": res["fp8"],
            "fp32_plain_mpix_s_incl_tr_rebake": round(pixels / plain_s / 1e6, 1)}


def c5(args):
    import mrpnn_amd as M
    ctx = M.Context()
    sizes = [(512, 512), (1280, 720), (1920, 1080)]
    rng = np.random.default_rng(5)
    res = {"config": "c5", "workload": "Poisson arrivals over 32 fBm volumes (128^3..512^3), "
                                       "sizes {512^2, 1280x720, 1920x1080}, random orbit / sun / alpha"}
    for name, prec, plain, rate, dur in (("bf16", M.BF16, False, args.serve_rate, args.serve_seconds),
                                         ("fp8", M.FP8, False, args.serve_rate, args.serve_seconds),
                                         ("bf16_at_300rps", M.BF16, False, 300.0, args.serve_seconds),
                                         ("fp32_plain", M.FP32, True, args.serve_rate, min(0.25, args.serve_seconds))):
        srv = M.Server(ctx, prec, plain=plain, max_batch_views=32, max_batch_pixels=8 << 20, max_inflight=2, g=G)
        ctx.set_option(M.OPT_TR_SLOTS, 128)
        ids = [srv.add_fbm_volume(int(rng.choice([128, 256, 384, 512])), 6, 4.0, 100 + k) for k in range(32)]
        # admission (fBm generation + pyramid) happens before the timed load: one tiny request per volume
        import torch
        tiny = torch.zeros((1, 4), device="cuda")
        for vid in ids:
            srv.submit(vid, M.orbit_view(1, 1, g=G), tiny)
        srv.drain()
        srv.poll(64)
        t0 = time.perf_counter()
        rep = srv.run_poisson(rate, dur, sizes, seed=7, buffers_per_size=48)
        st = srv.stats()
        _log("c5 %s: %s (%.1f s)" % (name, rep, time.perf_counter() - t0))
        res[name] = {"offered_rps": rate, "window_s": dur, "requests": rep["requests"],
                     "mpix_s": round(rep["mpix_s"], 1), "p50_ms": round(rep["p50_ms"], 2),
                     "p99_ms": round(rep["p99_ms"], 2), "max_ms": round(rep["max_ms"], 2),
                     "wall_s": round(rep["wall_s"], 2), "batches": st.batches,
                     "views_per_batch": round(st.batched_views / max(st.batches, 1), 2),
                     "admissions": st.admissions}
        srv.close()
    ctx.close()
    return res


def c5sweep(args):
    """C5's Poisson load at offered rates from 100 to 2000 requests/s (BF16; windows long enough
    for about 400 requests): Mpix/s, p50 and p99 latency per rate, the highest rate whose p99
    stays within 16 ms, and the per-request device cost that bounds the rate."""
    import torch

    import mrpnn_amd as M
    ctx = M.Context()
    ctx.set_option(M.OPT_TR_SLOTS, 128)
    sizes = [(512, 512), (1280, 720), (1920, 1080)]
    rng = np.random.default_rng(5)
    specs = [(int(rng.choice([128, 256, 384, 512])), 100 + k) for k in range(32)]
    srv = M.Server(ctx, M.BF16, max_batch_views=32, max_batch_pixels=8 << 20, max_inflight=2, g=G)
    ids = [srv.add_fbm_volume(N, 6, 4.0, seed) for N, seed in specs]
    tiny = torch.zeros((1, 4), device="cuda")
    for vid in ids:
        srv.submit(vid, M.orbit_view(1, 1, g=G), tiny)
    srv.drain()
    srv.poll(64)
    rows = []
    for rate in (100.0, 200.0, 300.0, 400.0, 500.0, 600.0, 700.0, 800.0, 1000.0, 1500.0, 2000.0):
        window = max(0.5, 400.0 / rate)
        rep = srv.run_poisson(rate, window, sizes, seed=int(rate), buffers_per_size=48)
        row = {"offered_rps": rate, "window_s": round(window, 2), "requests": rep["requests"],
               "served_rps": round(rep["completed"] / rep["wall_s"], 1), "mpix_s": round(rep["mpix_s"], 1),
               "p50_ms": round(rep["p50_ms"], 2), "p99_ms": round(rep["p99_ms"], 2), "max_ms": round(rep["max_ms"], 2)}
        _log("c5sweep: %s" % row)
        rows.append(row)
    srv.close()
    ok = [r["offered_rps"] for r in rows if r["p99_ms"] <= 16.0]
    # per-request device cost of the mix: one TR bake (a new sun) plus the network over its pixels
    vol = M.Volume(ctx, 512)
    vol.fill_fbm(6, 4.0, 1)
    vol.update()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    cost = {}
    for w, h in sizes:
        outs = torch.zeros((w * h, 4), device="cuda")
        ts = []
        for i in range(8):
            sun = (0.3 + 0.05 * i, 0.8, 0.2)
            v = M.orbit_view(w, h, dist=2.0, azimuth=0.5, elevation=0.2, fov_deg=40.0, light=sun, alpha=64.0, g=G,
                             albedo=0.99, seed=1)
            torch.cuda.synchronize()
            ev[0].record()
            ctx.prepare_views([vol], [v])
            ctx.render_batch([vol], [v], [outs], M.BF16)
            ev[1].record()
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        cost["%dx%d" % (w, h)] = round(float(np.median(ts[2:])), 3)
    vol.close()
    ctx.close()
    mean_ms = float(np.mean(list(cost.values())))
    return {"config": "c5sweep", "workload": "C5 Poisson load over 32 fBm volumes (128^3..512^3), sizes {512^2, "
                                             "1280x720, 1920x1080}, BF16, one sun per request",
            "rates": rows, "max_rate_p99_le_16ms": max(ok) if ok else None,
            "request_ms_with_bake": cost, "mix_mean_request_ms": round(mean_ms, 3),
            "device_bound_rps": round(1000.0 / mean_ms, 1)}


def tile(args):
    """Tile feature gather, the training-input path (MRPNNInputData.fetch_input at 1920x1080:
    a 240x135 grid of pixels with 64 sub-rays each over a 512^3 cloud): fetches per second,
    sub-ray marches per second, the oracle on the host cores over the same frame, and parity
    of every row against it."""
    import torch

    import mrpnn_amd as M
    import oracle as O
    ctx = M.Context()
    N = 512
    vol = M.Volume(ctx, N)
    vol.fill_fbm(6, 4.0, 1)
    vol.update()
    data = M.MRPNNInputData(1920, 1080, vol)
    rows = data.w * data.h
    res = {"config": "tile", "workload": "fetch_input at 1920x1080 (240x135 pixels x 64 sub-rays, 773 floats "
                                         "per pixel) over a 512^3 fBm cloud, default camera and light, "
                                         "all 64 sub-pixel offsets"}
    for alpha in (1.0, 4.0, 16.0):
        data.fetch_input(alpha=alpha)  # LUT and TR mips for this (g, light, alpha)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(64):
            data.fetch_input(alpha=alpha, pixelOffset=(k % 8, k // 8))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 64
        act = float((data.tensor[..., 0] != -1000.0).float().mean())
        res["alpha_%g" % alpha] = {"ms_per_fetch": round(dt * 1e3, 4), "fetches_s": round(1.0 / dt, 1),
                                   "rows_s": round(rows / dt, 1), "subrays_s": round(64 * rows / dt, 1),
                                   "active_frac": round(act, 4)}
        _log("tile alpha %g: %.3f ms per fetch, active %.3f" % (alpha, dt * 1e3, act))
    if args.no_cpu:
        ctx.close()
        return res
    # CPU baseline and parity: the oracle's GatherData restatement on the host cores, one frame
    alpha = 4.0
    got = data.fetch_input(alpha=alpha, pixelOffset=(3, 5)).reshape(rows, -1).cpu().numpy()
    threads = cpu_threads()
    O.set_threads(threads)
    sc = O.Scene.from_arrays(vol.get_data(), N, 1.0, vol.download_mips(), vol.download_tr(), ctx.download_hg_lut())
    v = M.camera_view(data.w, data.h, (0.67085, -0.03808, -0.04856), (0.34281, 0.70711, 0.61845), alpha=alpha, g=G)
    t0 = time.perf_counter()
    want, active = sc.tile_gather(O.View.from_buffer_copy(bytes(v)), 3, 5)
    cpu_s = time.perf_counter() - t0
    act = active > 0
    rel = np.abs(got[act, :581] - want[act, :581]) / np.maximum(np.abs(want[act, :581]), 1e-3)
    rel[:, 577:579] = np.abs(got[act, 577:579] - want[act, 577:579]) * (2e-4 / 1e-6)
    res["cpu_baseline"] = {"rows_s": round(rows / cpu_s, 1), "seconds": round(cpu_s, 3), "cores": threads, "cpu": cpu_info(),
                           "kind": "port", "sample": "one full 240x135 frame, alpha 4, offset (3, 5)"}
    res["parity_vs_oracle"] = {"active_equal": bool(np.array_equal(got[:, 0] != -1000.0, act)),
                               "t_grid_max_abs": float(np.abs(got[:, 581:] - want[:, 581:]).max()),
                               "rows_within_2e-4": round(float(np.mean(rel.max(axis=1) <= 2e-4)), 5)}
    ctx.close()
    return res


def interactive(args):
    """Interactive progressive renderer (the GUI path, VolumeRender::Render into caller-owned
    target / histogram / RGBA8 buffers) at 1024x1024 over a 512^3 cloud: frames per second of
    the MRPNN predictor (BF16 and FP32, with and without the checkerboard), of the
    delta-tracking path (512 bounces, albedo 1, the GUI defaults) with and without the
    histogram-similarity filter and of the RPNN predictor (BF16, with and without the
    checkerboard); the oracle's filter on the host cores over the same frame as
    baseline and check."""
    import torch

    import mrpnn_amd as M
    import oracle as O
    ctx = M.Context()
    N, W = 512, 1024
    vol = M.Volume(ctx, N)
    vol.fill_fbm(6, 4.0, 1)
    vol.update()
    it = M.Interactive(ctx, W, W)
    target, histo, rgba = it.buffers()
    v = M.camera_view(W, W, (0.67085, -0.03808, -0.04856), (0.34281, 0.70711, 0.61845), alpha=8.0, g=G)
    v.jitter = -1.0
    res = {"config": "interactive", "workload": "1024x1024 frames over a 512^3 fBm cloud (6 octaves), GUI camera "
                                                "and sun, alpha 8, ACES; PT 512 bounces, albedo 1"}
    cases = (("mrpnn_bf16", M.RT_MRPNN, M.BF16, 1, False, 32), ("mrpnn_bf16_checker", M.RT_MRPNN, M.BF16, 1, True, 32),
             ("mrpnn_fp32", M.RT_MRPNN, M.FP32, 1, False, 8), ("rpnn_bf16", M.RT_RPNN, M.BF16, 1, False, 8),
             ("rpnn_bf16_checker", M.RT_RPNN, M.BF16, 1, True, 8), ("pt_no_denoise", M.RT_PT, M.FP32, 512, False, 8),
             ("pt_denoise", M.RT_PT, M.FP32, 512, True, 8))  # the PT cases last: the filter check reads their state
    calls = 0  # the renderer's checkerboard phase flips on every call
    for name, rt, prec, mb, chk, frames in cases:
        it.set_checkerboard(chk)
        denoise = name == "pt_denoise"
        for f in range(2):
            it.render(vol, v, target, histo, rgba, rt, prec, mb, f, M.TONE_ACES, denoise)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(2, 2 + frames):
            it.render(vol, v, target, histo, rgba, rt, prec, mb, f, M.TONE_ACES, denoise)
        torch.cuda.synchronize()
        calls += 2 + frames
        last_frame, last_flip = 1 + frames, (calls - 1) % 2
        dt = (time.perf_counter() - t0) / frames
        res[name] = {"ms_per_frame": round(dt * 1e3, 3), "fps": round(1.0 / dt, 1)}
        _log("interactive %s: %.3f ms per frame" % (name, dt * 1e3))
    # the filter alone on the last PT state: the frame's denoise minus the plain resolve
    t = target.cpu().numpy()
    h = histo.cpu().numpy()
    res["pt_filter_ms"] = round(res["pt_denoise"]["ms_per_frame"] - res["pt_no_denoise"]["ms_per_frame"], 3)
    var = np.abs(h[:, 31] - h[:, 32] ** 2)
    res["filtered_pixel_frac"] = round(float(np.mean(var > 0.01)), 4)
    if args.no_cpu:
        ctx.close()
        return res
    threads = cpu_threads()
    O.set_threads(threads)
    got = rgba.cpu().numpy().view(np.uint32)
    t0 = time.perf_counter()
    want = O.denoise(t, h, W, W, last_frame, last_flip, 1, 1.0, M.TONE_ACES)
    cpu_s = time.perf_counter() - t0
    d = np.abs(got.view(np.uint8).reshape(-1, 4).astype(int) - want.view(np.uint8).reshape(-1, 4).astype(int))
    res["cpu_baseline"] = {"filter_ms_per_frame": round(cpu_s * 1e3, 1), "cores": threads, "cpu": cpu_info(), "kind": "port",
                           "sample": "the histogram filter and ACES over one 1024x1024 PT frame"}
    res["parity_vs_oracle"] = {"rgba_equal_frac": round(float(np.mean(d.max(axis=1) == 0)), 5),
                               "rgba_max_step": int(d.max())}
    ctx.close()
    return res


def ingest(args):
    """Volume ingest and environment lighting: parse rates of the loaders (a 256^3 .vox and a
    128^3 .txt written here), LoadData of a 512^3 .bin including upload and pyramid, the RGBE
    reader on a 2048x1024 RLE image, and the cost of the environment map on the 1080p BF16
    batched frame over a 512^3 cloud.  numpy's text parser is the CPU comparison."""
    import struct
    import tempfile

    import torch

    import mrpnn_amd as M
    res = {"config": "ingest", "workload": "256^3 .vox and 128^3 .txt parse, 512^3 .bin LoadData, 2048x1024 RGBE read, "
                                           "1080p BF16 frame with and without an environment map"}
    rng = np.random.default_rng(1)
    with tempfile.TemporaryDirectory() as d:
        vals = rng.integers(0, 256, 256 ** 3).astype(np.uint8)
        lut = [b"%d" % i for i in range(256)]
        lines = [b" ".join(lut[v] for v in vals[i:i + 64]) for i in range(0, vals.size, 64)]
        vox = os.path.join(d, "model.vox")
        with open(vox, "wb") as f:
            f.write(b"width 256\nheight 256\ndepth 256\n" + b"\n".join(lines) + b"\n")
        mb = os.path.getsize(vox) / 1e6
        t0 = time.perf_counter()
        g, n, md = M.read_volume(vox)
        dt = time.perf_counter() - t0
        assert np.array_equal(g, vals.astype(np.float32) * np.float32(64.0 / 255.0))
        res["vox_256"] = {"file_mb": round(mb, 1), "seconds": round(dt, 3), "mb_s": round(mb / dt, 1)}
        txt = os.path.join(d, "grid.txt")
        gt = rng.uniform(0, 1, 128 ** 3).astype(np.float32)
        with open(txt, "w") as f:
            f.write("128\n")
            np.savetxt(f, gt.reshape(-1, 8), fmt="%.9g")
        mb = os.path.getsize(txt) / 1e6
        t0 = time.perf_counter()
        g, n, md = M.read_volume(txt)
        dt = time.perf_counter() - t0
        assert np.array_equal(g, gt)
        t0 = time.perf_counter()
        gn = np.loadtxt(txt, skiprows=1, dtype=np.float32).ravel()
        dn = time.perf_counter() - t0
        assert np.array_equal(gn, gt)
        res["txt_128"] = {"file_mb": round(mb, 1), "seconds": round(dt, 3), "mb_s": round(mb / dt, 1),
                          "numpy_loadtxt_seconds": round(dn, 3)}
        ctx = M.Context()
        big = os.path.join(d, "CLOUD0")
        grid = rng.uniform(0, 1, 512 ** 3).astype(np.float32)
        with open(big + ".bin", "wb") as f:
            f.write(struct.pack("<i", 512))
            grid.tofile(f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vol = M.Volume.load(ctx, big)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        vol2 = M.Volume.load(ctx, big)  # the named dataset's host grid is cached
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res["bin_512_load"] = {"first_s": round(t1 - t0, 3), "cached_s": round(t2 - t1, 3),
                               "max_density": vol.max_density}
        vol2.close()
        vol.close()
        # RGBE: 2048x1024, RLE scanlines of a smooth sky
        W, H = 2048, 1024
        hdr = os.path.join(d, "sky.hdr")
        yy = np.linspace(0, 1, H)[:, None]
        xx = np.linspace(0, 1, W)[None, :]
        img = np.stack([0.5 + 0.5 * yy * np.ones_like(xx), 0.7 * np.ones((H, W)), 1.0 + 0 * xx * yy], -1)
        m = img.max(-1)
        e = np.ceil(np.log2(m)).astype(int)
        rgbe = np.concatenate([np.clip(img / np.ldexp(1.0, e)[..., None] * 256, 0, 255).astype(np.uint8),
                               (e + 128).astype(np.uint8)[..., None]], -1)
        with open(hdr, "wb") as f:
            f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (H, W))
            for j in range(H):
                f.write(bytes([2, 2, W >> 8, W & 255]))
                for c in range(4):
                    comp = rgbe[j, :, c]
                    i = 0
                    while i < W:
                        r = 1
                        while i + r < W and comp[i + r] == comp[i] and r < 127:
                            r += 1
                        if r >= 2:
                            f.write(bytes([128 + r, comp[i]]))
                            i += r
                        else:
                            f.write(bytes([1, comp[i]]))
                            i += 1
        t0 = time.perf_counter()
        env = M.read_hdr(hdr)
        res["hdr_2048x1024_s"] = round(time.perf_counter() - t0, 4)
        # the environment map on the batched 1080p frame
        vol = M.Volume(ctx, 512)
        vol.fill_fbm(6, 4.0, 1)
        vol.update()
        v = M.orbit_view(1920, 1080, dist=1.6, azimuth=0.5, elevation=0.3, fov_deg=45.0, light=_sun(40, 30),
                         alpha=64.0, g=G)
        out = torch.zeros((1920 * 1080, 4), device="cuda")
        for name, e_map in (("no_env", None), ("env", env)):
            ctx.set_env_map(e_map)
            ctx.prepare_views([vol], [v])
            ctx.render_batch([vol], [v], [out], M.BF16)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(20):
                ctx.render_batch([vol], [v], [out], M.BF16)
            torch.cuda.synchronize()
            res["frame_1080p_bf16_ms_" + name] = round((time.perf_counter() - t0) / 20 * 1e3, 3)
        ctx.close()
    return res


def rpnn(args):
    """RPNN, the paper's comparison predictor (RenderCamera<true, RPNN>), at 1920x1080 over a
    512^3 cloud with the headline camera and sun: frames per second of the FP32 path (the
    reference's arithmetic) and of the BF16 MFMA path, the BF16 error against FP32, the
    MRPNN BF16 batched frame of the same view for comparison, and the oracle's RPNN on the
    host cores over a 960x540 frame of the same camera as baseline and check."""
    import torch

    import mrpnn_amd as M
    import oracle as O
    ctx = M.Context()
    N, W, H = 512, 1920, 1080
    vol = M.Volume(ctx, N)
    vol.fill_fbm(6, 4.0, 1)
    vol.update()
    ctx.load_rpnn()
    sun = _sun(30.0, 40.0)

    def view(w, h):
        return M.orbit_view(w, h, dist=2.0, azimuth=0.6, elevation=0.25, fov_deg=40.0, light=sun, alpha=64.0, g=G,
                            albedo=0.99, seed=1, jitter=0.5)

    v = view(W, H)
    out = torch.zeros((W * H, 4), device="cuda")
    res = {"config": "rpnn", "workload": "RPNN frame at 1920x1080 over a 512^3 fBm cloud (6 octaves), orbit camera "
                                         "(distance 2, fov 40), alpha 64, one sun",
           "network_flop_per_pixel": 2 * 1292200}
    frames = {}
    for name, prec, reps in (("bf16", M.BF16, 10), ("fp32", M.FP32, 3)):
        vol.render_rpnn(v, prec, out)
        torch.cuda.synchronize()
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            vol.render_rpnn(v, prec, out)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
        frames[name] = out[:, 0].cpu().numpy().copy()
        active = int((out[:, 3] > 0).sum())
        dt = float(np.mean(lat))
        res[name] = {"ms_per_frame": round(dt * 1e3, 3), "mpix_s": round(W * H / dt / 1e6, 2),
                     "p99_ms": round(float(np.quantile(lat, 0.99)) * 1e3, 3), "active_pixels": active,
                     "network_tflops": round(active * res["network_flop_per_pixel"] / dt / 1e12, 1)}
        _log("rpnn %s: %.2f ms per frame, %d active pixels" % (name, dt * 1e3, active))
    res["speedup_bf16_vs_fp32"] = round(res["fp32"]["ms_per_frame"] / res["bf16"]["ms_per_frame"], 2)
    res["bf16_vs_fp32"] = _rel(frames["bf16"], frames["fp32"])
    # MRPNN on the same view (batched BF16 path)
    ctx.prepare_views([vol], [v])
    ctx.render_batch([vol], [v], [out], M.BF16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        ctx.prepare_views([vol], [v])
        ctx.render_batch([vol], [v], [out], M.BF16)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    res["mrpnn_bf16_same_view"] = {"ms_per_frame": round(dt * 1e3, 3)}
    res["rpnn_bf16_over_mrpnn_bf16_time"] = round(res["bf16"]["ms_per_frame"] / (dt * 1e3), 2)
    if args.no_cpu:
        ctx.close()
        return res
    # the oracle on the host cores: a 960x540 frame of the same camera, and parity against it
    threads = cpu_threads()
    O.set_threads(threads)
    sc = O.Scene.from_arrays(vol.get_data(), N, 1.0, vol.download_mips(), vol.download_tr(), ctx.download_hg_lut())
    vs = view(960, 540)
    t0 = time.perf_counter()
    ref = O.render_rpnn(sc, O.View.from_buffer_copy(bytes(vs)))
    cpu_s = time.perf_counter() - t0
    act = ref["active"] > 0
    got32 = vol.render_rpnn(vs, M.FP32).cpu().numpy()
    got16 = vol.render_rpnn(vs, M.BF16).cpu().numpy()
    rel = np.abs(got32[act, 0] - ref["out"][act, 0]) / np.maximum(ref["out"][act, 0], 1e-6)
    res["cpu_baseline"] = {"pixels_s": round(960 * 540 / cpu_s, 1), "active_pixels_s": round(int(act.sum()) / cpu_s, 1),
                           "seconds": round(cpu_s, 3), "cores": threads, "cpu": cpu_info(), "kind": "port",
                           "sample": "one 960x540 frame of the same camera (descriptor and FP32 network)"}
    res["parity_vs_oracle"] = {"fp32_median_rel": float(np.median(rel)), "fp32_max_rel": float(rel.max()),
                               "sky_equal": bool(np.array_equal(got32[~act], ref["out"][~act])),
                               "bf16": _rel(got16[act, 0], ref["out"][act, 0])}
    ctx.close()
    return res
