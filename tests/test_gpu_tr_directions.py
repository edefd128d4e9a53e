"""GPU parity of the TR bake for lights along every principal axis and sign.

The level-0 bake stages its density windows along the light's principal axis A (one kernel
instantiation per axis) and walks its slab range towards the exit face (up or down the axis), so
each (axis, sign) is its own code path: every one is compared with the oracle's bake (same
tolerance as the single-direction parity test: device vs libm expf, 2e-5 relative), and so is a
light with three equal components (the tie picks axis x)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIGHTS = [
    (-0.90, 0.30, 0.20),   # x, down
    (0.30, 0.90, 0.20),    # y, up
    (0.10, -0.95, 0.30),   # y, down
    (0.20, 0.30, 0.93),    # z, up
    (-0.20, 0.10, -0.97),  # z, down
    (0.577, 0.577, 0.577),  # tie: x
]


@pytest.fixture(scope="module")
def scene(ctx, orc):
    import mrpnn_amd as M
    N = 64
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, 3)
    vol.update()
    sc = orc.Scene(orc.fbm_volume(N, 5, 4.0, 3), N)
    return vol, sc


@pytest.mark.parametrize("light", LIGHTS, ids=["x-", "y+", "y-", "z+", "z-", "xyz"])
def test_tr_mips_per_axis(scene, light):
    vol, sc = scene
    L = np.array(light, np.float32)
    L /= np.linalg.norm(L)
    vol.update_tr(L, 64.0)
    sc.update_tr(L, 64.0)
    np.testing.assert_allclose(vol.download_tr(), sc.tr, rtol=2e-5, atol=1e-7)
