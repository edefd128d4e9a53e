"""Serving loop on the GPU: admission by an HBM budget with least-recently-used eviction,
continuous batching, results identical alone and in the mix, the plain baseline mode and a
short open-loop Poisson load.

Bar: bit-exact between a request served in a mixed batch (with volumes evicted and admitted
while it waits) and the same request rendered alone on a fresh volume.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 0.857
FBM = [(32, 1), (48, 2), (64, 3), (40, 4)]


@pytest.fixture(scope="module")
def sctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mrpnn_amd as M
    c = M.Context()
    yield c
    c.close()


def _requests(n, seed=5):
    import mrpnn_amd as M
    rng = np.random.default_rng(seed)
    sizes = [(40, 32), (24, 24), (33, 17), (64, 48), (1, 1)]
    out = []
    for i in range(n):
        w, h = sizes[i % len(sizes)]
        sun = rng.normal(size=3)
        sun[1] = abs(sun[1]) + 0.2
        v = M.orbit_view(w, h, azimuth=float(rng.uniform(0, 6.28)), elevation=float(rng.uniform(-0.2, 0.5)),
                         light=tuple(sun), alpha=float(rng.uniform(16, 128)), g=G, view_id=i)
        out.append((int(rng.integers(0, len(FBM))), v))
    return out


def _alone(ctx, vid, view):
    """The request rendered by itself on a freshly generated volume."""
    import torch
    import mrpnn_amd as M
    N, seed = FBM[vid]
    vol = M.Volume(ctx, N)
    vol.fill_fbm(5, 4.0, seed)
    vol.update()
    v = M.View.from_buffer_copy(bytes(view))
    v.volume_index = 0
    out = torch.zeros((v.height * v.width, 4), device="cuda")
    ctx.prepare_views([vol], [v])
    ctx.render_batch([vol], [v], [out], M.BF16)
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    vol.close()
    return res


def test_mix_with_eviction_equals_alone(sctx):
    import torch
    import mrpnn_amd as M
    per_volume = M.lib().mrp_volume_bytes(64)
    srv = M.Server(sctx, M.BF16, max_batch_views=5, max_inflight=2, g=G,
                   hbm_limit_bytes=int(2.5 * per_volume))  # two volumes resident at a time
    ids = [srv.add_fbm_volume(N, 5, 4.0, seed) for N, seed in FBM]
    reqs = _requests(14)
    outs = []
    for i, (vid, v) in enumerate(reqs):
        o = torch.full((v.height * v.width, 4), -3.0, device="cuda")
        outs.append(o)
        srv.submit(ids[vid], v, o, tag=i)
        if i % 3 == 2:
            srv.step()  # arrivals interleaved with scheduling
    srv.drain()
    torch.cuda.synchronize()
    done = srv.poll()
    assert sorted(c.user_tag for c in done) == list(range(len(reqs)))
    st = srv.stats()
    assert st.completed == len(reqs) and st.pending == 0
    assert st.resident_volumes <= 2 and st.evictions >= 2 and st.admissions == st.evictions + st.resident_volumes
    assert st.batches < len(reqs)  # requests were batched
    for i, (vid, v) in enumerate(reqs):
        assert np.array_equal(outs[i].cpu().numpy(), _alone(sctx, vid, v)), i
    srv.close()


def test_plain_mode_matches_plain_render(sctx):
    import torch
    import mrpnn_amd as M
    srv = M.Server(sctx, M.FP32, plain=True, g=G)
    ids = [srv.add_fbm_volume(N, 5, 4.0, seed) for N, seed in FBM[:2]]
    reqs = _requests(4, seed=9)
    outs = []
    for i, (vid, v) in enumerate(reqs):
        vid = vid % 2
        o = torch.zeros((v.height * v.width, 4), device="cuda")
        outs.append((vid, v, o))
        srv.submit(ids[vid], v, o, tag=i)
    srv.drain()
    torch.cuda.synchronize()
    for vid, v, o in outs:
        N, seed = FBM[vid]
        vol = M.Volume(sctx, N)
        vol.fill_fbm(5, 4.0, seed)
        vol.update()
        assert np.array_equal(o.cpu().numpy(), vol.render_plain(v))
        vol.close()
    srv.close()


def test_volume_larger_than_budget_is_an_error(sctx):
    import torch
    import mrpnn_amd as M
    srv = M.Server(sctx, M.BF16, g=G, hbm_limit_bytes=1 << 20)
    vid = srv.add_fbm_volume(32, 5, 4.0, 1)
    v = _requests(1)[0][1]
    o = torch.zeros((v.height * v.width, 4), device="cuda")
    srv.submit(vid, v, o)
    with pytest.raises(M.MrpError):
        srv.drain()
    srv.close()


def test_poisson_load_completes(sctx):
    import mrpnn_amd as M
    srv = M.Server(sctx, M.BF16, g=G)
    for N, seed in FBM:
        srv.add_fbm_volume(N, 5, 4.0, seed)
    rep = srv.run_poisson(400.0, 0.25, [(64, 64), (96, 54)], seed=3, buffers_per_size=8)
    assert rep["requests"] > 20 and rep["completed"] == rep["requests"]
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    srv.close()
