"""GPU parity of the RPNN comparison predictor (core/radiancePredict.cu:567-722) against the
oracle: the network's input rows (same march, light frame and mip samples; the basis and
gamma through device vs libm sinf / cosf / acosf: 1e-4 relative), the FP32 network on
identical rows (the reference's summation order on both sides: 1e-5 relative), and whole
FP32 frames including sky pixels and composition."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rscene(ctx, orc):
    import mrpnn_amd as M
    N = 48
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, 2)
    vol.update()
    sc = orc.Scene(orc.fbm_volume(N, 5, 4.0, 2), N)
    ctx.load_rpnn()
    return vol, sc


def _view(orc, M, w=28, h=22):
    v = orc.orbit_view(w, h, alpha=32.0, light=(0.3, 0.9, 0.2), azimuth=0.7, elevation=0.25)
    return v, M.View.from_buffer_copy(bytes(v))


def test_rpnn_rows(ctx, orc, rscene):
    import mrpnn_amd as M
    vol, sc = rscene
    v, vm = _view(orc, M)
    rows, out = vol.rpnn_features(vm)
    ref = orc.render_rpnn(sc, v)
    act = ref["active"] > 0
    assert act.sum() > 100 and (~act).sum() > 50
    got = rows.cpu().numpy()
    assert np.all(got[~act] == 0)
    # the basis' device vs libm sinf / cosf move far samples by ulps of their position
    np.testing.assert_allclose(got[act, :2250], ref["rows"][act, :2250], rtol=1e-4, atol=2e-6)
    d = np.abs(got[act, :2250] - ref["rows"][act, :2250])
    assert np.median(d[d > 0] / np.abs(ref["rows"][act, :2250])[d > 0]) < 1e-6  # ulp-sized where they differ
    np.testing.assert_allclose(got[act, 2250], ref["rows"][act, 2250], rtol=0, atol=2e-6)


def test_rpnn_fp32_network_on_identical_rows(ctx, orc, rscene):
    import torch
    vol, sc = rscene
    import mrpnn_amd as M
    v, _ = _view(orc, M)
    ref = orc.render_rpnn(sc, v)
    rows = ref["rows"][ref["active"] > 0]
    want = orc.rpnn_predict_f32(rows)
    got = ctx.rpnn_predict_rows(torch.tensor(rows, device="cuda")).cpu().numpy()
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    print("rpnn fp32 vs oracle: max rel %.2e, equal %.4f" % (rel.max(), np.mean(got == want)))
    assert rel.max() <= 1e-5


def test_rpnn_fp32_frame(ctx, orc, rscene):
    import mrpnn_amd as M
    vol, sc = rscene
    v, vm = _view(orc, M)
    got = vol.render_rpnn(vm, M.FP32).cpu().numpy()
    ref = orc.render_rpnn(sc, v)
    act = ref["active"] > 0
    assert np.all(got[~act] == 0)
    # inputs differ by ulps (basis sinf / cosf); the recurrent network carries them to the output
    rel = np.abs(got[act, 0] - ref["out"][act, 0]) / np.maximum(ref["out"][act, 0], 1e-6)
    print("rpnn fp32 frame vs oracle: median rel %.2e max %.2e" % (np.median(rel), rel.max()))
    assert np.median(rel) < 1e-5 and rel.max() < 1e-3
    np.testing.assert_allclose(got[act, 3], ref["out"][act, 3], rtol=1e-6)


def test_rpnn_bf16_frame_close_to_fp32(ctx, orc, rscene):
    """The BF16 MFMA network (fused, descriptor fragments from HBM) against the FP32 path on the
    same frame: identical sky pixels and distances, radiance within a few percent on average
    (ten recurrent levels of bf16 activations and weights)."""
    import mrpnn_amd as M
    vol, sc = rscene
    v, vm = _view(orc, M, 40, 32)
    a = vol.render_rpnn(vm, M.FP32).cpu().numpy()
    b = vol.render_rpnn(vm, M.BF16).cpu().numpy()
    sky = a[:, 3] == 0
    assert np.array_equal(a[sky], b[sky]) and np.array_equal(a[:, 3], b[:, 3])
    m = a[:, 0] > 1e-3 * a[:, 0].max()
    rel = np.abs(b[m, 0] - a[m, 0]) / a[m, 0]
    print("rpnn bf16 vs fp32: mean rel %.4f p99 %.4f over %d pixels" % (rel.mean(), np.quantile(rel, 0.99), m.sum()))
    assert rel.mean() < 0.05


def test_rpnn_bf16_network_on_identical_rows(ctx, orc, rscene):
    """The BF16 MFMA network alone (fragments packed from the oracle's rows) against the
    oracle's FP32 network."""
    import torch
    import mrpnn_amd as M
    vol, sc = rscene
    v, _ = _view(orc, M)
    ref = orc.render_rpnn(sc, v)
    rows = ref["rows"][ref["active"] > 0]
    want = orc.rpnn_predict_f32(rows)
    got = ctx.rpnn_predict_rows(torch.tensor(rows, device="cuda"), M.BF16).cpu().numpy()
    m = want > 1e-3 * want.max()
    rel = np.abs(got[m] - want[m]) / want[m]
    print("rpnn bf16 rows vs oracle fp32: mean rel %.4f p99 %.4f" % (rel.mean(), np.quantile(rel, 0.99)))
    assert rel.mean() < 0.05 and np.all(np.isfinite(got)) and np.all(got >= 0)


def test_rpnn_bf16_descriptor_matches_rows(ctx, orc, rscene):
    """The BF16 render's own descriptor (LDS boxes, fused lattice coordinates) against the exact
    FP32 rows of the same view fed to the same BF16 network: the two inputs differ by float
    rounding only, so after bf16 rounding the outputs agree to a small fraction of a percent."""
    import torch
    import mrpnn_amd as M
    vol, sc = rscene
    for w, h, azi in ((40, 32, 0.7), (36, 30, 2.3)):
        v = orc.orbit_view(w, h, alpha=32.0, light=(-0.5, 0.4, 0.77), azimuth=azi, elevation=0.4)
        vm = M.View.from_buffer_copy(bytes(v))
        rows, _ = vol.rpnn_features(vm)
        got4 = vol.render_rpnn(vm, M.BF16).cpu().numpy()
        act = got4[:, 3] > 0  # distance: zero for sky pixels
        assert act.sum() > 300
        idx = torch.from_numpy(np.flatnonzero(act)).cuda()
        y = ctx.rpnn_predict_rows(rows[idx].contiguous(), M.BF16)
        want = (y.cpu().numpy() * np.float32(v.light_color[0])) * np.float32(v.scatter_rate[0])
        got = got4[act, 0]
        m = want > 1e-3 * want.max()
        rel = np.abs(got[m] - want[m]) / want[m]
        print("rpnn bf16 render vs bf16(rows): median %.2e p99 %.2e max %.2e" % (np.median(rel), np.quantile(rel, 0.99), rel.max()))
        assert np.median(rel) < 1e-3 and np.quantile(rel, 0.99) < 1e-2


def test_rpnn_bf16_passes_and_tails(ctx, orc, rscene):
    """The BF16 path in many small passes (4,096 pixels: tail tiles of 16 and 128 pixels in
    every pass) gives the same frame as one pass: pixels are independent of the chunking."""
    import mrpnn_amd as M
    vol, sc = rscene
    v = orc.orbit_view(160, 128, alpha=32.0, light=(0.3, 0.9, 0.2), azimuth=0.7, elevation=0.25)
    vm = M.View.from_buffer_copy(bytes(v))
    one = vol.render_rpnn(vm, M.BF16).cpu().numpy()
    active = int((one[:, 3] > 0).sum())
    assert active > 2 * 4096 and active % 128 != 0
    default = ctx.get_option(M.OPT_RPNN_CHUNK)
    try:
        ctx.set_option(M.OPT_RPNN_CHUNK, 4096)
        many = vol.render_rpnn(vm, M.BF16).cpu().numpy()
    finally:
        ctx.set_option(M.OPT_RPNN_CHUNK, default)
    np.testing.assert_array_equal(one, many)


def test_rpnn_empty_volume_and_bad_arguments(ctx, orc):
    """An empty volume has no active pixel: the frame is the sky (distance 0), the same in both
    precisions; unknown precisions and chunk sizes are refused."""
    import mrpnn_amd as M
    vol = M.Volume(ctx, 32)
    vol.update()
    v = orc.orbit_view(24, 20, alpha=32.0)
    vm = M.View.from_buffer_copy(bytes(v))
    a = vol.render_rpnn(vm, M.FP32).cpu().numpy()
    b = vol.render_rpnn(vm, M.BF16).cpu().numpy()
    assert np.all(a[:, 3] == 0) and np.array_equal(a, b)
    with pytest.raises(RuntimeError):
        vol.render_rpnn(vm, 7)
    for bad in (100, 4095, 2 ** 21):
        with pytest.raises(RuntimeError):
            ctx.set_option(M.OPT_RPNN_CHUNK, bad)
    vol.close()
