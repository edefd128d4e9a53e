"""GPU side of volume ingest and environment lighting: loaded volumes render as the same grid
set directly, the named-dataset cache serves a second load without the file, and with an
environment map the sky pixels of the plain, batched and interactive predictors and the
escaping paths of the path tracer see SkyBox (core/render.cuh:24-27) as the oracle computes
it (device vs libm atan2f / acosf: 5e-5 relative on a map with steep texel steps)."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(N, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(N) + 0.5) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    g = np.clip(1.0 - r / 0.4, 0, None) * (0.7 + 0.3 * rng.uniform(size=r.shape))
    return g.astype(np.float32).ravel()


def _view(orc, M, w, h, **kw):
    v = orc.orbit_view(w, h, fov_deg=40.0, alpha=24.0, **kw)
    return v, M.View.from_buffer_copy(bytes(v))


def test_loaded_txt_renders_as_set_data(ctx, orc, tmp_path):
    import mrpnn_amd as M
    N = 32
    g = _grid(N, 1)
    p = tmp_path / "ball.txt"
    with open(p, "w") as f:
        f.write("%d\n" % N)
        for i in range(0, N ** 3, 8):
            f.write(" ".join("%.9g" % v for v in g[i:i + 8]) + "\n")
    loaded = M.Volume.load(ctx, str(p))
    assert loaded.max_density == g.max()
    ref = M.Volume(ctx, N)
    ref.set_data(g, float(g.max()))
    ref.update()
    v, vm = _view(orc, M, 24, 20)
    assert np.array_equal(loaded.render_plain(vm), ref.render_plain(vm))
    loaded.close()
    ref.close()


def test_named_dataset_cache_serves_without_file(ctx, orc, tmp_path):
    import os

    import mrpnn_amd as M
    N = 16
    g = _grid(N, 2)
    base = tmp_path / "MODEL1"
    with open(str(base) + ".bin", "wb") as f:
        f.write(struct.pack("<i", N) + g.tobytes())
    a = M.Volume.load(ctx, str(base))
    os.remove(str(base) + ".bin")
    b = M.Volume.load(ctx, str(base))
    assert a.max_density == b.max_density == np.float32(23.8960018)
    v, vm = _view(orc, M, 16, 16)
    assert np.array_equal(a.render_plain(vm), b.render_plain(vm))
    a.close()
    b.close()


@pytest.fixture()
def env_scene(ctx, orc):
    import mrpnn_amd as M
    N = 32
    g = _grid(N, 3)
    vol = M.Volume(ctx, N)
    vol.set_data(g, float(g.max()))
    vol.update()
    sc = orc.Scene(g, N, float(g.max()))
    rng = np.random.default_rng(4)
    env = rng.uniform(0.0, 2.0, (16, 32, 4)).astype(np.float32)
    yield vol, sc, env
    ctx.set_env_map(None)
    ctx.set_env_exp(1.0)
    orc.set_env(None)
    vol.close()


def test_sky_pixels_see_the_environment(ctx, orc, env_scene):
    import mrpnn_amd as M
    import torch
    vol, sc, env = env_scene
    v, vm = _view(orc, M, 40, 32, light=(0.3, 0.9, 0.2))
    plain0 = vol.render_plain(vm)
    ctx.set_env_map(env)
    ctx.set_env_exp(1.3)
    orc.set_env(env, 1.3)
    plain = vol.render_plain(vm)
    L = np.asarray(v.light_dir[:], np.float32)
    sc.update_tr(L, v.alpha)
    ctx.update_hg_lut(v.g, 1)
    sc.set_lut(ctx.download_hg_lut())
    ref = sc.render_predict(v)
    sky = ref["active"] == 0
    assert sky.sum() > 100 and (~sky).sum() > 100
    # a texel-to-texel step of the random map times an ulp of atan2f / acosf: a few 1e-6
    np.testing.assert_allclose(plain[sky], ref["out"][sky], rtol=5e-5, atol=1e-6)
    assert np.all(plain[sky, :3].max(axis=1) > 0)
    assert np.array_equal(plain[~sky], plain0[~sky])  # the cloud itself does not change
    out = torch.zeros((v.width * v.height, 4), device="cuda")
    ctx.prepare_views([vol], [vm])
    ctx.render_batch([vol], [vm], [out], M.BF16)
    assert np.array_equal(out.cpu().numpy()[sky], plain[sky])


def test_escaping_paths_see_the_environment(ctx, orc, env_scene):
    import mrpnn_amd as M
    vol, sc, env = env_scene
    ctx.set_env_map(env)
    ctx.set_env_exp(0.8)
    orc.set_env(env, 0.8)
    v, vm = _view(orc, M, 16, 12, light=(0.3, 0.9, 0.2))
    L = np.asarray(v.light_dir[:], np.float32)
    vol.update_tr(L, v.alpha)
    sc.update_tr(L, v.alpha)
    mean, _ = vol.pt_render(vm, 8, 16, nee_baked=1)
    mo, _, _ = sc.pt_render(v, 8, 16, nee_baked=1)
    mg = mean.cpu().numpy()
    close = np.abs(mg - mo) <= 1e-5 * np.maximum(np.abs(mo), 1e-3)
    assert close.mean() > 0.9
    assert np.mean(mg.sum(axis=1) > 0) > 0.95  # sky rays see the map at once, the others on escape


def test_interactive_sky_uses_the_environment(ctx, orc, env_scene):
    import mrpnn_amd as M
    vol, sc, env = env_scene
    ctx.set_env_map(env)
    W, H = 24, 20
    v, vm = _view(orc, M, W, H, light=(0.3, 0.9, 0.2), jitter=-1.0, frame=2)
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(False)
    target, histo, rgba = it.buffers()
    it.render(vol, vm, target, histo, rgba, M.RT_MRPNN, M.FP32, randseed=2, tone=M.TONE_NONE)
    plain = vol.render_plain(vm).reshape(H, W, 4)[::-1].reshape(-1, 4)
    assert np.array_equal(target.cpu().numpy(), np.maximum(plain[:, :3], 0))
    it.close()
