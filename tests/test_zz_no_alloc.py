"""mrp_render_batch_noalloc, the batched predictor with its per-call device buffers in the
library's static scratch array, against mrp_render_batch through the C ABI.

Bar: bit-exact.  Both entry points run the same march and network kernels on the same view
and volume tables, and each pixel's result does not depend on where its record sits in the
active list.  Two calls in a row (the second smaller, so it runs over the first one's stale
records and tables), every precision; batches beyond the stated limits are refused with
MRP_ERR_INVALID; free device memory does not move across a call.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 0.857
SUNS = [(0.866, 0.5, 0.0), (0.2, 0.9, 0.3), (-0.5, 0.4, 0.7)]
MAX_VIEWS = 64
MAX_RECORDS = 2097152


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mrpnn_amd as M
    c = M.Context()
    vols = []
    for N, seed in ((64, 3), (48, 2)):
        v = M.Volume(c, N)
        v.fill_fbm(5, 4.0, seed)
        v.update()
        vols.append(v)
    yield M, c, vols
    c.close()


def _views(M, spec):
    return [M.orbit_view(w, h, azimuth=az, elevation=0.3, light=SUNS[si], alpha=alpha, g=G, view_id=i, volume_index=vi)
            for i, (w, h, vi, si, az, alpha) in enumerate(spec)]


# first call: five views (two table chunks of the staging kernel), one of them 1920x1080, in
# all 2,078,976 of the 2,097,152 records; second call: two small views
FIRST = [(1920, 1080, 0, 0, 0.3, 64.0), (64, 48, 1, 1, 1.0, 32.0), (33, 17, 0, 2, 2.0, 96.0), (1, 1, 1, 0, 0.5, 64.0),
         (40, 32, 0, 1, 4.0, 64.0)]
SECOND = [(24, 24, 1, 2, 1.5, 48.0), (16, 40, 0, 0, 3.0, 128.0)]


def _call(M, c, vols, views, prec, noalloc):
    import torch
    outs = [torch.full((v.height * v.width, 4), -7.0 if noalloc else -3.0, device="cuda") for v in views]
    vh = (ctypes.c_void_p * len(vols))(*[v.h for v in vols])
    va = (M.View * len(views))(*views)
    oa = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
    fn = M.lib().mrp_render_batch_noalloc if noalloc else M.lib().mrp_render_batch
    rc = fn(c.h, vh, len(vols), va, len(views), int(prec), oa)
    assert rc == 0, M.lib().mrp_last_error()
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("prec", ["bf16", "fp8", "fp32"])
def test_noalloc_equals_render_batch(env, prec):
    M, c, vols = env
    p = M.PRECISIONS[prec]
    for spec in (FIRST, SECOND):
        views = _views(M, spec)
        c.prepare_views(vols, views)
        ref = _call(M, c, vols, views, p, False)
        got = _call(M, c, vols, views, p, True)
        for i in range(len(views)):
            assert np.array_equal(ref[i], got[i]), "%s view %d of %d differs" % (prec, i, len(views))
        assert sum(int((r[:, 3] > 0).sum()) for r in ref) > 1000 or spec is SECOND


def test_noalloc_keeps_free_memory(env):
    import torch
    M, c, vols = env
    views = _views(M, FIRST)
    c.prepare_views(vols, views)
    _call(M, c, vols, views, M.BF16, True)
    outs = [torch.zeros((v.height * v.width, 4), device="cuda") for v in views]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    c.render_batch_noalloc(vols, views, outs, M.BF16)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_refuses_beyond_limits(env):
    import torch
    M, c, vols = env
    lib = M.lib()
    vh = (ctypes.c_void_p * len(vols))(*[v.h for v in vols])
    out = torch.zeros(16, device="cuda")
    # 2048x1040 needs 256 * 130 tiles * 64 = 2,129,920 records
    big = _views(M, [(2048, 1040, 0, 0, 0.3, 64.0)])
    assert 64 * ((2048 + 7) // 8) * ((1040 + 7) // 8) > MAX_RECORDS
    c.prepare_views(vols, big)
    va = (M.View * 1)(*big)
    oa = (ctypes.c_void_p * 1)(out.data_ptr())
    assert lib.mrp_render_batch_noalloc(c.h, vh, len(vols), va, 1, M.BF16, oa) == 1
    assert b"limit" in lib.mrp_last_error()
    many = _views(M, [(8, 8, 0, 0, 0.1 * i, 64.0) for i in range(MAX_VIEWS + 1)])
    c.prepare_views(vols, many[:1])
    for v in many:
        v.ori[:] = many[0].ori[:]
    va = (M.View * len(many))(*many)
    oa = (ctypes.c_void_p * len(many))(*([out.data_ptr()] * len(many)))
    assert lib.mrp_render_batch_noalloc(c.h, vh, len(vols), va, len(many), M.BF16, oa) == 1
    assert b"n_views" in lib.mrp_last_error()
    torch.cuda.synchronize()
