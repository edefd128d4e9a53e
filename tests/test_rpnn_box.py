"""The BF16 RPNN descriptor kernel (k_rpnn_desc) stages, per pixel and level, a 5 x 5 x 5 texel
box of the level's mip whose origin is the floor of the lattice's lowest coordinate, and clamps
every sample's box coordinate into [0, 3.99].  That is exact only if every lattice point lies
less than 3.99 texels above the box origin on each axis (the clamp then only absorbs rounding at
the lower face).  The 225-point lattice (x, y in [-2, 2], z in [-2, 6] in the light frame
eX = normalize(cross(LX, X)), eY = normalize(cross(eX, LX)), LX) has spacing 1/16, 1/8, then
1/4 texel for levels 0, 1, 2..9 (2^(l-12) of the unit box on mip max(l - 2, 0) of 256 >> mip
texels), so per axis its extent is at most sp * (4|eX_i| + 4|eY_i| + 8|LX_i|) <= sqrt(6) = 2.45
texels.  Checked here in float32 over random frames, including near-degenerate ones."""
import numpy as np


def _norm(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_lattice_fits_the_box():
    rng = np.random.default_rng(7)
    n = 20000
    X = _norm(rng.normal(size=(n, 3))).astype(np.float32)
    LX = _norm(rng.normal(size=(n, 3))).astype(np.float32)
    # a few nearly parallel / antiparallel shading and light directions
    LX[:200] = _norm(X[:200] + 1e-3 * rng.normal(size=(200, 3))).astype(np.float32)
    LX[200:400] = _norm(-X[200:400] + 1e-3 * rng.normal(size=(200, 3))).astype(np.float32)
    eX = _norm(np.cross(LX, X)).astype(np.float32)
    eY = _norm(np.cross(eX, LX)).astype(np.float32)
    pos = rng.uniform(-0.5, 0.5, size=(n, 3)).astype(np.float32)
    g = np.arange(-2, 3, dtype=np.float32)
    z = np.arange(-2, 7, dtype=np.float32)
    xs, ys, zs = (a.ravel() for a in np.meshgrid(g, g, z, indexing="ij"))  # 225 points
    worst = 0.0
    for l in range(10):
        mip = max(l - 2, 0)
        res = 256 >> mip
        sp = np.float32(2.0 ** (l - 12) * res)  # lattice spacing in texels
        u0 = (pos + np.float32(0.5)) * np.float32(res) - np.float32(0.5)
        a, b, c = eX * sp, eY * sp, LX * sp
        lo = np.floor(u0 - 2 * np.abs(a) - 2 * np.abs(b) + np.minimum(-2 * c, 6 * c))
        # every lattice point, box coordinates (float32 fused form as the kernel)
        d = (u0 - lo)[:, None, :] + xs[None, :, None] * a[:, None, :] + ys[None, :, None] * b[:, None, :] \
            + zs[None, :, None] * c[:, None, :]
        assert d.min() > -1e-3, (l, d.min())  # at most rounding below the lower face
        worst = max(worst, float(d.max()))
        assert d.max() < 3.5, (l, d.max())
    assert worst > 2.0  # the bound is not vacuous: some frames use most of the box
