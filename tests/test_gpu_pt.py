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
