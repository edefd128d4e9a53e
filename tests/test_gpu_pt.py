"""GPU path tracer against the oracle: same Philox streams, so paths coincide except where a
device-vs-libm ulp flips a collision decision; FP64 accumulators merge exactly across chunks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pt_scene(ctx, orc):
    import mrpnn_amd as M
    N = 48
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, 3)
    vol.update()
    L = np.array([0.3, 0.9, 0.2], np.float32)
    L /= np.linalg.norm(L)
    vol.update_tr(L, 32.0)
    sc = orc.Scene(orc.fbm_volume(N, 5, 4.0, 3), N)
    sc.update_tr(L, 32.0)
    return vol, sc, L


def test_pt_matches_oracle(orc, pt_scene):
    import mrpnn_amd as M
    vol, sc, L = pt_scene
    v = orc.orbit_view(12, 10, light=L, alpha=32.0)
    vm = M.View.from_buffer_copy(bytes(v))
    mean, m2 = vol.pt_render(vm, 24, 48)
    mo, m2o, _ = sc.pt_render(v, 24, 48)
    mg = mean.cpu().numpy()
    close = np.abs(mg - mo) <= 1e-5 * np.maximum(np.abs(mo), 1e-3)
    assert close.mean() > 0.9
    # pixels whose paths diverged still agree statistically
    se = np.sqrt((m2o + m2.cpu().numpy()) / (24 * 23) + 1e-12)
    assert np.all(np.abs(mg - mo) <= 6 * se + 1e-6)


def test_pt_chunked_accumulation_is_exact(orc, pt_scene):
    import mrpnn_amd as M
    vol, sc, L = pt_scene
    v = orc.orbit_view(10, 8, light=L, alpha=32.0)
    vm = M.View.from_buffer_copy(bytes(v))
    a, a2 = vol.pt_render(vm, 12, 32)
    b, b2 = vol.pt_render(vm, 12, 32, chunk=5)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()) and np.array_equal(a2.cpu().numpy(), b2.cpu().numpy())


def test_pt_baked_nee_mode(orc, pt_scene):
    import mrpnn_amd as M
    vol, sc, L = pt_scene
    v = orc.orbit_view(8, 8, light=L, alpha=32.0)
    vm = M.View.from_buffer_copy(bytes(v))
    mean, _ = vol.pt_render(vm, 16, 32, nee_baked=1)
    mo, _, _ = sc.pt_render(v, 16, 32, nee_baked=1)
    close = np.abs(mean.cpu().numpy() - mo) <= 1e-5 * np.maximum(np.abs(mo), 1e-3)
    assert close.mean() > 0.9


def _tasks(n, seed=0):
    rng = np.random.default_rng(seed)
    ori = rng.uniform(-0.35, 0.35, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    L = rng.normal(size=(n, 3))
    L[:, 1] = np.abs(L[:, 1])
    L /= np.linalg.norm(L, axis=1, keepdims=True)
    alpha = rng.uniform(8, 48, n).astype(np.float32)
    g = rng.uniform(0.0, 0.9, n).astype(np.float32)
    sc = rng.uniform(0.9, 0.999, n).astype(np.float32)
    return ori, d.astype(np.float32), L.astype(np.float32), alpha, g, sc


def test_point_radiance_matches_oracle(orc, pt_scene):
    """GetSamples on the GPU against the oracle: same Philox streams per (task, sample)."""
    vol, sc, L = pt_scene
    n, spp = 40, 48
    ori, d, Ls, alpha, g, scat = _tasks(n)
    mg, m2g = vol.point_radiance(ori, d, Ls, alpha, g, scat, (1.0, 0.9, 0.8), 64, spp, seed=7)
    mo, m2o = sc.point_radiance(ori, d, Ls, alpha, g, scat, (1.0, 0.9, 0.8), 64, spp, 7)
    mg = mg.cpu().numpy().astype(np.float64)
    close = np.abs(mg - mo) <= 1e-5 * np.maximum(np.abs(mo), 1e-3)
    assert close.mean() > 0.9
    se = np.sqrt((m2o + m2g.cpu().numpy()) / (spp * (spp - 1)) + 1e-12)
    assert np.all(np.abs(mg - mo) <= 6 * se + 1e-6)


def test_tr_samples_match_oracle(orc, pt_scene):
    vol, sc, L = pt_scene
    ori, d, _, _, _, _ = _tasks(64, 1)
    got = vol.tr_samples(ori, d, L, 32.0, 0.7, 32, seed=3).cpu().numpy()
    want = sc.tr_samples(ori, d, L, 32.0, 0.7, 32, 3)
    close = np.abs(got - want) <= 1e-5 * np.maximum(np.abs(want), 1e-4)
    assert close.mean() > 0.9 and np.all(np.isfinite(got))


def test_cached_bake_and_pt_match_oracle(orc, pt_scene):
    """Fill_BakeTR at 24^3 against the oracle texel by texel, then the path tracer with NEE
    from that bake (nee 2, the reference's PT) against the oracle with the same bake."""
    import mrpnn_amd as M
    vol, sc, L = pt_scene
    res = 24
    vol.bake_cached_tr(L, 32.0, 1.0, seed=5, res=res)
    got = vol.download_cached_tr(res)
    want = sc.bake_cached_tr(L, 32.0, 1.0, 5, res, np.arange(res ** 3))
    close = np.abs(got - want) <= 1e-5 * np.maximum(np.abs(want), 1e-4)
    assert close.mean() > 0.97 and np.all((got >= 0) & (got <= 1))
    v = orc.orbit_view(10, 8, light=L, alpha=32.0)
    vm = M.View.from_buffer_copy(bytes(v))
    mean, m2 = vol.pt_render(vm, 16, 32, nee_baked=2)
    mo, m2o = sc.pt_render_list_cached(v, np.arange(80), 16, 32, got, res)
    mg = mean.cpu().numpy()
    close = np.abs(mg - mo) <= 1e-5 * np.maximum(np.abs(mo), 1e-3)
    assert close.mean() > 0.9
    # refused: another alpha, another tr_scale, another light than the bake
    for kw in ({"alpha": 16.0}, {"light": (0.0, 1.0, 0.0)}):
        v2 = orc.orbit_view(4, 4, light=kw.get("light", L), alpha=kw.get("alpha", 32.0))
        with pytest.raises(M.MrpError):
            vol.pt_render(M.View.from_buffer_copy(bytes(v2)), 1, 4, nee_baked=2)
    with pytest.raises(M.MrpError):
        vol.pt_render(vm, 1, 4, nee_baked=2, tr_scale=2.0)


def test_cached_bake_default_resolution(orc, pt_scene):
    """The reference's 768^3 bake (res 0): sampled texels equal the oracle's."""
    vol, sc, L = pt_scene
    vol.bake_cached_tr(L, 32.0, 1.0, seed=2, res=0)
    got = vol.download_cached_tr(768)
    idx = np.random.default_rng(4).integers(0, 768 ** 3, 4000)
    want = sc.bake_cached_tr(L, 32.0, 1.0, 2, 768, idx)
    close = np.abs(got[idx] - want) <= 1e-5 * np.maximum(np.abs(want), 1e-4)
    assert close.mean() > 0.97
    del got
