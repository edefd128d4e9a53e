"""Batched path on the GPU: several resident volumes, views under different suns, ragged
image sizes, the TR cache, HIP-graph replay and the launch-shape autotuner.

Bar: bit-exact.  Every 16-record network tile holds pixels of one 8x8 screen tile of one view
(padded compaction) and the FP8 activation scales are fixed per layer, so a view's pixels are
computed by the same instructions on the same data whether it is rendered alone or in a
mix, eagerly or from a graph, with any persistent grid size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 0.857


@pytest.fixture(scope="module")
def bctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mrpnn_amd as M
    c = M.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def volumes(bctx):
    import mrpnn_amd as M
    vols = []
    for N, seed in ((32, 1), (48, 2), (64, 3), (40, 4)):
        v = M.Volume(bctx, N)
        v.fill_fbm(5, 4.0, seed)
        v.update()
        vols.append(v)
    empty = M.Volume(bctx, 16)  # all-zero grid: every ray misses
    empty.update()
    vols.append(empty)
    return vols


def _views():
    import mrpnn_amd as M
    suns = [(0.866, 0.5, 0.0), (0.2, 0.9, 0.3), (-0.5, 0.4, 0.7), (0.1, 0.3, -0.9)]
    spec = [  # (w, h, volume, sun, azimuth, alpha)
        (40, 32, 0, 0, 0.0, 64.0),
        (24, 24, 1, 1, 1.0, 32.0),
        (33, 17, 2, 2, 2.0, 96.0),   # ragged: not a multiple of the 8x8 tile
        (1, 1, 2, 3, 0.5, 64.0),     # one pixel
        (64, 48, 2, 0, 4.0, 64.0),   # same volume as views 2 and 3, third sun
        (16, 40, 3, 1, 3.0, 128.0),
        (20, 12, 4, 2, 0.0, 64.0),   # empty volume
    ]
    out = []
    for i, (w, h, vi, si, az, alpha) in enumerate(spec):
        out.append(M.orbit_view(w, h, azimuth=az, elevation=0.3, light=suns[si], alpha=alpha, g=G, view_id=i,
                                volume_index=vi))
    return out


def _render(ctx, vols, views, prec):
    import torch
    outs = [torch.full((v.height * v.width, 4), -7.0, device="cuda") for v in views]
    ctx.prepare_views(vols, views)
    ctx.render_batch(vols, views, outs, prec)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("prec", ["bf16", "fp8", "fp32"])
def test_mixed_batch_equals_alone(bctx, volumes, prec):
    import mrpnn_amd as M
    views = _views()
    mix = _render(bctx, volumes, views, M.PRECISIONS[prec])
    for i, v in enumerate(views):
        alone = _render(bctx, volumes, [v], M.PRECISIONS[prec])[0]
        assert np.array_equal(alone, mix[i]), "view %d differs alone vs in the mix" % i
    for i, img in enumerate(mix):
        assert np.all(np.isfinite(img)) and np.all(img >= 0), i
    assert np.all(mix[-1] == 0)  # empty volume: no active pixel, zeros written


def test_mixed_batch_fp32_equals_plain(bctx, volumes):
    """The batched FP32 path equals the plain per-frame kernel view by view (each with the
    TR mips of its own sun)."""
    import mrpnn_amd as M
    views = _views()
    mix = _render(bctx, volumes, views, M.FP32)
    for i, v in enumerate(views):
        plain = volumes[v.volume_index].render_plain(v)
        assert np.array_equal(plain, mix[i]), i


def test_tr_cache_capacity_and_eviction(bctx, volumes):
    import mrpnn_amd as M
    views = _views()  # 7 distinct (volume, sun, alpha) keys
    ref = _render(bctx, volumes, views, M.BF16)
    bctx.set_option(M.OPT_TR_SLOTS, 3)
    try:
        with pytest.raises(M.MrpError):
            bctx.prepare_views(volumes, views)
        # batches that fit evict older sets and re-bake them later with identical results
        for a, b in ((0, 3), (3, 6), (0, 3), (4, 7)):
            got = _render(bctx, volumes, views[a:b], M.BF16)
            for k, img in enumerate(got):
                assert np.array_equal(img, ref[a + k]), (a, k)
    finally:
        bctx.set_option(M.OPT_TR_SLOTS, 64)


def test_graph_replay_matches_eager(bctx, volumes):
    """Graph mode: each replay (with eager work and synchronisations between replays) equals
    the eager result; new views of the same launch shape replay the same graph with the new
    staged parameters."""
    import torch
    import mrpnn_amd as M
    views = _views()
    eager = _render(bctx, volumes, views, M.BF16)
    views2 = _views()
    for v in views2:  # same shapes, other camera and sun
        w = M.orbit_view(v.width, v.height, azimuth=1.7, elevation=0.1, light=(0.3, 0.8, 0.5), alpha=v.alpha,
                         g=G, view_id=v.view_id, volume_index=v.volume_index)
        v.ori[:], v.up[:], v.right[:], v.light_dir[:] = w.ori[:], w.up[:], w.right[:], w.light_dir[:]
    eager2 = _render(bctx, volumes, views2, M.BF16)
    bctx.set_option(M.OPT_GRAPH, 1)
    try:
        for rep in range(4):
            got = _render(bctx, volumes, views if rep % 2 == 0 else views2, M.BF16)
            want = eager if rep % 2 == 0 else eager2
            for i in range(len(views)):
                assert np.array_equal(got[i], want[i]), (rep, i)
            x = torch.randn(1 << 20, device="cuda")  # eager work between replays
            (x * 2).sum().item()
    finally:
        bctx.set_option(M.OPT_GRAPH, 0)


def test_autotune_keeps_results(bctx, volumes):
    import torch
    import mrpnn_amd as M
    views = _views()
    ref = _render(bctx, volumes, views, M.BF16)
    outs = [torch.zeros((v.height * v.width, 4), device="cuda") for v in views]
    bctx.prepare_views(volumes, views)
    w, ms = bctx.autotune(volumes, views, outs, M.BF16)
    assert w in (4, 8, 12, 16, 24, 32) and ms > 0
    lds = bctx.get_option(M.OPT_LDS_WEIGHTS)
    assert lds in (0, 1)
    if not lds:
        assert bctx.get_option(M.OPT_MFMA_WAVES_PER_CU) == w
    assert bctx.get_option(M.OPT_TILE_W) in (4, 8, 16, 32, 64)
    got = _render(bctx, volumes, views, M.BF16)
    for i in range(len(views)):
        assert np.array_equal(got[i], ref[i]), i
    bctx.set_option(M.OPT_MFMA_WAVES_PER_CU, 12)
    bctx.set_option(M.OPT_LDS_WEIGHTS, 1)
    bctx.set_option(M.OPT_TILE_W, 8)


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
def test_tile_shapes_give_identical_images(bctx, volumes, prec):
    """March tiles of 4x16, 16x4, 32x2 and 64x1 pixels give the 8x8 tiles' images bit for bit
    (each pixel's march, gather and network do not depend on its tile): eager, graph replay and
    the no-allocation entry point."""
    import torch
    import mrpnn_amd as M
    views = _views()
    p = M.PRECISIONS[prec]
    assert bctx.get_option(M.OPT_TILE_W) == 8
    ref = _render(bctx, volumes, views, p)
    try:
        for tw in (4, 16, 32, 64):
            bctx.set_option(M.OPT_TILE_W, tw)
            assert bctx.get_option(M.OPT_TILE_W) == tw
            got = _render(bctx, volumes, views, p)
            for i in range(len(views)):
                assert np.array_equal(got[i], ref[i]), (tw, i)
            bctx.set_option(M.OPT_GRAPH, 1)
            got = _render(bctx, volumes, views, p)
            bctx.set_option(M.OPT_GRAPH, 0)
            for i in range(len(views)):
                assert np.array_equal(got[i], ref[i]), ("graph", tw, i)
            outs = [torch.full((v.height * v.width, 4), -7.0, device="cuda") for v in views]
            bctx.render_batch_noalloc(volumes, views, outs, p)
            torch.cuda.synchronize()
            for i in range(len(views)):
                assert np.array_equal(outs[i].cpu().numpy(), ref[i]), ("noalloc", tw, i)
        for bad in (0, 2, 12, 128):
            with pytest.raises(M.MrpError):
                bctx.set_option(M.OPT_TILE_W, bad)
    finally:
        bctx.set_option(M.OPT_GRAPH, 0)
        bctx.set_option(M.OPT_TILE_W, 8)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["bf16", "fp8", "bf16_1t"])
def test_lds_weights_equal_l2_weights(bctx, volumes, prec):
    """The LDS-resident-weight kernel (one 768-thread workgroup per CU) and the L1/L2-weight
    kernel evaluate the same fragments in the same order: identical images."""
    import mrpnn_amd as M
    views = _views()
    p = M.PRECISIONS[prec]
    try:
        bctx.set_option(M.OPT_LDS_WEIGHTS, 1)
        a = _render(bctx, volumes, views, p)
        bctx.set_option(M.OPT_LDS_WEIGHTS, 0)
        for w in (4, 12):
            bctx.set_option(M.OPT_MFMA_WAVES_PER_CU, w)
            b = _render(bctx, volumes, views, p)
            for i in range(len(views)):
                assert np.array_equal(a[i], b[i]), (prec, w, i)
    finally:
        bctx.set_option(M.OPT_LDS_WEIGHTS, 1)
        bctx.set_option(M.OPT_MFMA_WAVES_PER_CU, 12)
