"""GPU parity of the interactive progressive path (mrp_render_interactive) against the oracle,
frame by frame: the frame's radiance against the oracle's path tracer (PT) or the plain
predictor (MRPNN), then the blend, the histograms, the checkerboard fill, the filter and the
tone map against the oracle fed with the same radiance.

Tolerances: target and histograms bit-exact (the same float operations in the same order);
the 8-bit output within one step per channel on at most 1% of the pixels (device and libm
powf in the tone maps differ by ulps); PT radiance as the path-tracer tests (same Philox
streams, > 90% of the pixels within 1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALPHA = 32.0


@pytest.fixture(scope="module")
def iscene(ctx, orc):
    import mrpnn_amd as M
    N = 48
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, 3)
    vol.update()
    sc = orc.Scene(orc.fbm_volume(N, 5, 4.0, 3), N)
    L = np.array([0.3, 0.9, 0.2], np.float32)
    L /= np.linalg.norm(L)
    sc.update_tr(L, ALPHA)
    return vol, sc, L


def _views(orc, M, W, H, L, frame):
    v = orc.orbit_view(W, H, light=L, alpha=ALPHA, jitter=-1.0, seed=5, frame=frame, azimuth=0.4, elevation=0.2)
    return v, M.View.from_buffer_copy(bytes(v))


def _flip_rows(img, W, H):
    return np.ascontiguousarray(img.reshape(H, W, -1)[::-1].reshape(W * H, -1))


def _rgba_close(got, want):
    a = got.view(np.uint8).reshape(-1, 4).astype(int)
    b = want.view(np.uint8).reshape(-1, 4).astype(int)
    d = np.abs(a - b)
    assert d.max() <= 1 and np.mean(d.max(axis=1) > 0) <= 0.01, (d.max(), np.mean(d.max(axis=1) > 0))


@pytest.mark.parametrize("checker", [0, 1])
def test_pt_frames(ctx, orc, iscene, checker):
    import mrpnn_amd as M
    vol, sc, L = iscene
    W, H = 40, 32
    n = W * H
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(checker)
    it.set_exposure(1.2)
    target, histo, rgba = it.buffers()
    tgt = np.zeros((n, 3), np.float32)
    hst = np.zeros((n, 33), np.float32)
    for f in range(6):
        print("interactive PT checker %d frame %d" % (checker, f), flush=True)
        flip = f % 2
        v, vm = _views(orc, M, W, H, L, f)
        it.render(vol, vm, target, histo, rgba, M.RT_PT, max_bounce=8, randseed=f, tone=M.TONE_ACES, denoise=True)
        rad = it.radiance().cpu().numpy()
        r, c = np.divmod(np.arange(n), W)
        s = (((r + flip) % 2) + c) % 2 == 0 if checker else np.ones(n, bool)
        assert np.all(rad[~s, 3] == -2.0) and np.all(rad[s, 3] == 0.0)
        # the frame's paths: the oracle's path tracer on the same streams (render rows flipped)
        mo, _, _ = sc.pt_render(v, 1, 8, nee_baked=1)
        mo = _flip_rows(mo.astype(np.float32), W, H)
        close = np.abs(rad[s, :3] - mo[s]) <= 1e-4 * np.maximum(np.abs(mo[s]), 1e-3)
        assert close.mean() > 0.9
        # blend, histogram, denoise, tone map on the device radiance
        fnum = f // 2 if checker else f
        orc.pt_accumulate(tgt, hst, rad, W, H, checker, flip, fnum)
        assert np.array_equal(target.cpu().numpy(), tgt)
        assert np.array_equal(histo.cpu().numpy(), hst)
        want = orc.denoise(tgt, hst, W, H, f, flip, 1, 1.2, 2)
        _rgba_close(rgba.cpu().numpy().view(np.uint32), want)
    it.close()


@pytest.mark.parametrize("checker", [0, 1])
def test_mrpnn_frames_fp32(ctx, orc, iscene, checker):
    import mrpnn_amd as M
    vol, sc, L = iscene
    W, H = 36, 28
    n = W * H
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(checker)
    target, histo, rgba = it.buffers()
    tgt = np.zeros((n, 3), np.float32)
    r, c = np.divmod(np.arange(n), W)
    for f in range(4):
        print("interactive MRPNN checker %d frame %d" % (checker, f), flush=True)
        flip = f % 2
        v, vm = _views(orc, M, W, H, L, f)
        it.render(vol, vm, target, histo, rgba, M.RT_MRPNN, M.FP32, randseed=f, tone=M.TONE_GAMMA)
        rad = it.radiance().cpu().numpy()
        s = (((r + flip) % 2) + c) % 2 == 0 if checker else np.ones(n, bool)
        plain = _flip_rows(vol.render_plain(vm), W, H)
        assert np.array_equal(rad[s, :3], np.maximum(plain[s, :3], 0.0))
        assert np.array_equal(rad[s, 3], plain[s, 3]) and np.all(rad[~s, 3] == -2.0)
        orc.predict_commit(tgt, rad, W, H, checker, flip)
        hst = np.zeros((n, 33), np.float32)
        want = orc.reprojection_denoise(tgt, hst, 1.0, 1)
        assert np.array_equal(target.cpu().numpy(), tgt)
        h = histo.cpu().numpy()
        assert np.array_equal(h[:, 30:], hst[:, 30:])
        _rgba_close(rgba.cpu().numpy().view(np.uint32), want)
    it.close()


def test_mrpnn_bf16_frame_is_batched_render(ctx, orc, iscene):
    import mrpnn_amd as M
    import torch
    vol, sc, L = iscene
    W, H = 32, 24
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(False)
    target, histo, rgba = it.buffers()
    v, vm = _views(orc, M, W, H, L, 3)
    it.render(vol, vm, target, histo, rgba, M.RT_MRPNN, M.BF16, randseed=3, tone=M.TONE_ACES)
    rad = it.radiance().cpu().numpy()
    out = torch.zeros((W * H, 4), device="cuda")
    ctx.prepare_views([vol], [vm])
    ctx.render_batch([vol], [vm], [out], M.BF16)
    ref = _flip_rows(out.cpu().numpy(), W, H)
    assert np.array_equal(rad[:, :3], np.maximum(ref[:, :3], 0.0))
    it.close()


@pytest.mark.parametrize("checker", [0, 1])
def test_rpnn_frames(ctx, orc, iscene, checker):
    """RenderCamera<true, RPNN> in the interactive renderer: the frame's radiance is the RPNN
    frame of the same view (FP32: the reference's arithmetic; BF16: its MFMA network) on the
    frame's checkerboard half, then the same commit, histograms and tone map as MRPNN."""
    import mrpnn_amd as M
    vol, sc, L = iscene
    W, H = 36, 28
    n = W * H
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(checker)
    target, histo, rgba = it.buffers()
    tgt = np.zeros((n, 3), np.float32)
    r, c = np.divmod(np.arange(n), W)
    for f, prec in enumerate((M.FP32, M.FP32, M.BF16, M.BF16)):
        print("interactive RPNN checker %d frame %d" % (checker, f), flush=True)
        flip = f % 2
        v, vm = _views(orc, M, W, H, L, f)
        it.render(vol, vm, target, histo, rgba, M.RT_RPNN, prec, randseed=f, tone=M.TONE_GAMMA)
        rad = it.radiance().cpu().numpy()
        s = (((r + flip) % 2) + c) % 2 == 0 if checker else np.ones(n, bool)
        frame = _flip_rows(vol.render_rpnn(vm, prec).cpu().numpy(), W, H)
        assert np.array_equal(rad[s, :3], np.maximum(frame[s, :3], 0.0))
        assert np.array_equal(rad[s, 3], frame[s, 3]) and np.all(rad[~s, 3] == -2.0)
        assert (rad[s, 3] > 0).sum() > 50
        orc.predict_commit(tgt, rad, W, H, checker, flip)
        hst = np.zeros((n, 33), np.float32)
        want = orc.reprojection_denoise(tgt, hst, 1.0, 1)
        assert np.array_equal(target.cpu().numpy(), tgt)
        _rgba_close(rgba.cpu().numpy().view(np.uint32), want)
    it.close()


def test_rpnn_frames_refuse_other_precisions(ctx, orc, iscene):
    """RPNN runs in FP32 or BF16 only: the interactive renderer refuses FP8 and single-term BF16
    as mrp_render_rpnn does, before touching the caller's buffers."""
    import mrpnn_amd as M
    vol, sc, L = iscene
    W, H = 16, 12
    it = M.Interactive(ctx, W, H)
    target, histo, rgba = it.buffers()
    target.fill_(3.0)
    v, vm = _views(orc, M, W, H, L, 0)
    for prec in (M.FP8, M.BF16_1T):
        with pytest.raises(M.MrpError, match="RPNN"):
            it.render(vol, vm, target, histo, rgba, M.RT_RPNN, prec, randseed=0, tone=M.TONE_GAMMA)
        with pytest.raises(M.MrpError):
            vol.render_rpnn(vm, prec)
    assert bool((target == 3.0).all())
    it.close()


def test_switch_from_pt_clears_histograms(ctx, orc, iscene):
    import mrpnn_amd as M
    vol, sc, L = iscene
    W, H = 24, 16
    it = M.Interactive(ctx, W, H)
    it.set_checkerboard(False)
    target, histo, rgba = it.buffers()
    v, vm = _views(orc, M, W, H, L, 0)
    it.render(vol, vm, target, histo, rgba, M.RT_PT, max_bounce=4, randseed=0, tone=M.TONE_NONE)
    assert histo.cpu().numpy()[:, 30].min() == 1.0
    it.render(vol, vm, target, histo, rgba, M.RT_MRPNN, M.FP32, randseed=1, tone=M.TONE_NONE)
    h = histo.cpu().numpy()
    t = target.cpu().numpy()
    assert np.all(h[:, :30] == 0) and np.array_equal(h[:, 30], t[:, 0]) and np.array_equal(h[:, 32], t[:, 1])
    it.close()


def test_interactive_rejects_bad_arguments(ctx, iscene, orc):
    import mrpnn_amd as M
    vol, sc, L = iscene
    it = M.Interactive(ctx, 16, 16)
    target, histo, rgba = it.buffers()
    v, vm = _views(orc, M, 16, 16, L, 0)
    with pytest.raises(M.MrpError):
        it.render(vol, vm, target, histo, rgba, 9)  # unknown render type
    v2, vm2 = _views(orc, M, 16, 8, L, 0)
    with pytest.raises(M.MrpError):
        it.render(vol, vm2, target, histo, rgba, M.RT_MRPNN)
    with pytest.raises(M.MrpError):
        it.render(vol, vm, target, histo, rgba, M.RT_MRPNN, tone=7)
    it.close()
