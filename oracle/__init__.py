"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product library (mrpnn_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
WEIGHTS = os.path.join(HERE, "..", "mrpnn_amd", "assets", "mrpnn_weights.bin")
RPNN_WEIGHTS = os.path.join(HERE, "..", "mrpnn_amd", "assets", "rpnn_weights.bin")
RPNN_ROW = 2251  # 10 levels x 225 density samples, then gamma

NFEAT = 577
TILE_ROW = 773  # tile feature gather row: 577 features + 4 gamma extras + 192 t-grid values
PC = 192
MIP_TOTAL = sum((256 >> m) ** 3 for m in range(9))
TR_TOTAL = sum((128 >> m) ** 3 for m in range(8))
LUT = 512

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_u = ctypes.POINTER(ctypes.c_uint32)


class View(ctypes.Structure):
    """Same layout as mrp_view (include/mrp_c_api.h)."""
    _fields_ = [("ori", ctypes.c_float * 3), ("up", ctypes.c_float * 3), ("right", ctypes.c_float * 3),
                ("light_dir", ctypes.c_float * 3), ("light_color", ctypes.c_float * 3),
                ("scatter_rate", ctypes.c_float * 3), ("alpha", ctypes.c_float), ("g", ctypes.c_float),
                ("tan_half_fov", ctypes.c_float), ("jitter", ctypes.c_float), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("seed", ctypes.c_uint32), ("frame", ctypes.c_uint32),
                ("view_id", ctypes.c_uint32), ("volume_index", ctypes.c_int32)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_load_weights.argtypes = [_f, ctypes.c_int]
        L.orc_philox.argtypes = [_u, ctypes.c_uint32, ctypes.c_uint32, _u]
        L.orc_exp_f32.argtypes = [_f, ctypes.c_int, _f]
        L.orc_u01.argtypes = [ctypes.c_uint32]
        L.orc_u01.restype = ctypes.c_float
        L.orc_halton.argtypes = [ctypes.c_uint, ctypes.c_uint]
        L.orc_halton.restype = ctypes.c_float
        L.orc_stencil.argtypes = [ctypes.c_int, _f, _i, _i, _i]
        L.orc_predict_f32.argtypes = [_f, ctypes.c_int, _f, ctypes.c_float, _f]
        L.orc_predict_f64.argtypes = [_d, ctypes.c_int, _d, ctypes.c_double, _d]
        L.orc_predict_amax_f64.argtypes = [_d, ctypes.c_int, _d, ctypes.c_double, _d]
        L.orc_build_pyramid.argtypes = [_f, ctypes.c_int, _f]
        L.orc_bake_tr.argtypes = [_f, _f, ctypes.c_float, _f]
        L.orc_hg_lut_qmc.argtypes = [ctypes.c_float, ctypes.c_uint32, _f]
        L.orc_hg_ball_mean.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_hg_ball_mean.restype = ctypes.c_double
        L.orc_sample_lut.argtypes = [_f, ctypes.c_float, ctypes.c_float]
        L.orc_sample_lut.restype = ctypes.c_float
        vol = [_f, ctypes.c_int, ctypes.c_float, _f, _f, _f]
        L.orc_render_predict.argtypes = vol + [ctypes.POINTER(View), _f, _f, _i, _i, _f]
        L.orc_render_predict_rows.argtypes = vol + [ctypes.POINTER(View), ctypes.c_int, ctypes.c_int, _f]
        L.orc_render_predict_list.argtypes = vol + [ctypes.POINTER(View), _i, ctypes.c_int, _f, _f]
        L.orc_march.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, _f, ctypes.c_float, _f, _i, _i]
        L.orc_tile_gather.argtypes = vol + [ctypes.POINTER(View), ctypes.c_int, ctypes.c_int, _f, _i]
        L.orc_sample_t.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, _f, ctypes.c_float, _f]
        L.orc_gather_state.argtypes = vol + [ctypes.c_int, _f, _f, _f, _f, _f, ctypes.c_float, ctypes.c_float, _f]
        L.orc_basis.argtypes = [_f, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f]
        L.orc_pt_render_list.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, ctypes.POINTER(View), _i, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _d, _d]
        L.orc_pt_render.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, ctypes.POINTER(View), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float, _d, _d,
                                    ctypes.POINTER(ctypes.c_int64)]
        L.orc_pt_render_list_cached.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, ctypes.c_int, ctypes.POINTER(View),
                                                _i, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _d, _d]
        L.orc_bake_cached_tr.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.c_int, _f]
        L.orc_bake_cached_tr.restype = None
        L.orc_point_radiance.argtypes = [_f, ctypes.c_int, ctypes.c_float, ctypes.c_int, _f, _f, _f, _f, _f, _f, _f,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_float, _d, _d]
        L.orc_point_radiance.restype = None
        L.orc_tr_samples.argtypes = [_f, ctypes.c_int, ctypes.c_float, ctypes.c_int, _f, _f, _f, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_uint32, ctypes.c_float, _d]
        L.orc_tr_samples.restype = None
        L.orc_tr_stochastic.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, _f, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_uint32, ctypes.c_int, _f]
        L.orc_fbm_volume.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint32, _f]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _v = ctypes.c_void_p
        L.orc_rpnn_load.argtypes = [_f, ctypes.c_int]
        L.orc_rpnn_predict_f32.argtypes = [_f, ctypes.c_int, _f]
        L.orc_rpnn_predict_f64.argtypes = [_f, ctypes.c_int, _d]
        L.orc_rpnn_descriptor.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, _f, _f, _f, ctypes.c_float, _f, _f]
        L.orc_rpnn_descriptor.restype = None
        L.orc_render_rpnn.argtypes = [_f, ctypes.c_int, ctypes.c_float, _f, ctypes.POINTER(View), _f, _i, _f]
        L.orc_set_env.argtypes = [_f, ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_set_env.restype = None
        L.orc_sky_box.argtypes = [_f, _f]
        L.orc_sky_box.restype = None
        L.orc_pt_accumulate.argtypes = [_f, _v, _f, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int]
        L.orc_predict_commit.argtypes = [_f, _f, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_denoise.argtypes = [_f, _v, _u, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_int]
        L.orc_reprojection_denoise.argtypes = [_f, _v, _u, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.orc_tonemap.argtypes = [_f, ctypes.c_int, ctypes.c_float, ctypes.c_int, _u, _f]
        L.orc_histo_compare.argtypes = [_v, _v]
        L.orc_histo_compare.restype = ctypes.c_float
        w = np.fromfile(WEIGHTS, dtype="<f4")
        if L.orc_load_weights(w.ctypes.data_as(_f), int(w.size)) != 0:
            raise RuntimeError("oracle: weight asset does not match the network layout")
        _lib = L
    return _lib


def _p(a, t=_f):
    return None if a is None else a.ctypes.data_as(t)


def set_threads(n):
    return lib().orc_set_threads(int(n))


def weights():
    return np.fromfile(WEIGHTS, dtype="<f4")


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c, _u), int(key[0]), int(key[1]), _p(out, _u))
    return out


def u01(x):
    return lib().orc_u01(int(x))


def exp_f32(x):
    """The FP32 networks' exp (mrpnn_amd/csrc/fexp.h) on an array."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_exp_f32(_p(x), x.size, _p(out))
    return out


def halton(dim, index):
    return lib().orc_halton(dim, index)


def stencil(index):
    off = ctypes.c_float()
    mip, typ, loc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib().orc_stencil(index, ctypes.byref(off), ctypes.byref(mip), ctypes.byref(typ), ctypes.byref(loc))
    return off.value, mip.value, typ.value, loc.value


def predict_f32(feats, sr, g):
    feats = np.ascontiguousarray(feats, np.float32).reshape(-1, NFEAT)
    sr = np.asarray(sr, np.float32)
    out = np.zeros((feats.shape[0], 3), np.float32)
    lib().orc_predict_f32(_p(feats), feats.shape[0], _p(sr), float(g), _p(out))
    return out


def predict_f64(feats, sr, g):
    feats = np.ascontiguousarray(feats, np.float64).reshape(-1, NFEAT)
    sr = np.asarray(sr, np.float64)
    out = np.zeros((feats.shape[0], 3), np.float64)
    lib().orc_predict_f64(_p(feats, _d), feats.shape[0], _p(sr, _d), float(g), _p(out, _d))
    return out


def predict_amax_f64(feats, sr, g):
    """Per MFMA layer (the library's 60 FP8 layers, in order) the largest |input| of the FP64
    network over the rows."""
    feats = np.ascontiguousarray(feats, np.float64).reshape(-1, NFEAT)
    sr = np.asarray(sr, np.float64)
    out = np.zeros(60, np.float64)
    lib().orc_predict_amax_f64(_p(feats, _d), feats.shape[0], _p(sr, _d), float(g), _p(out, _d))
    return out


def fbm_volume(N, octaves, base_freq, seed):
    out = np.zeros(N ** 3, np.float32)
    lib().orc_fbm_volume(N, octaves, base_freq, seed, _p(out))
    return out


def build_pyramid(grid, N):
    grid = np.ascontiguousarray(grid, np.float32)
    out = np.zeros(MIP_TOTAL, np.float32)
    lib().orc_build_pyramid(_p(grid), N, _p(out))
    return out


def split_mips(mips):
    res, o = [], 0
    for m in range(9):
        r = 256 >> m
        res.append(mips[o:o + r ** 3].reshape(r, r, r))
        o += r ** 3
    return res


def split_tr(tr):
    res, o = [], 0
    for m in range(8):
        r = 128 >> m
        res.append(tr[o:o + r ** 3].reshape(r, r, r))
        o += r ** 3
    return res


def bake_tr(mips, light, alpha):
    mips = np.ascontiguousarray(mips, np.float32)
    light = np.asarray(light, np.float32)
    out = np.zeros(TR_TOTAL, np.float32)
    lib().orc_bake_tr(_p(mips), _p(light), float(alpha), _p(out))
    return out


def hg_lut(g, seed):
    out = np.zeros(LUT * LUT, np.float32)
    lib().orc_hg_lut_qmc(float(g), int(seed), _p(out))
    return out


def hg_ball_mean(costheta, angle, g, n_psi=96, n_phi=96):
    return lib().orc_hg_ball_mean(costheta, angle, g, n_psi, n_phi)


def sample_lut(lut, u, v):
    lut = np.ascontiguousarray(lut, np.float32)
    return lib().orc_sample_lut(_p(lut), u, v)


class Scene:
    """A volume with its pyramids on the host (the oracle's view of VolumeRender)."""

    def __init__(self, grid, N, max_density=1.0):
        self.grid = np.ascontiguousarray(grid, np.float32)
        self.N = N
        self.max_density = float(max_density)
        self.mips = build_pyramid(self.grid, N)
        self.tr = None
        self.lut = None

    def update_tr(self, light, alpha):
        self.tr = bake_tr(self.mips, light, alpha)

    def set_lut(self, lut):
        self.lut = np.ascontiguousarray(lut, np.float32)

    def _vol(self):
        return [_p(self.grid), self.N, self.max_density, _p(self.mips), _p(self.tr), _p(self.lut)]

    def render_predict(self, view, want_feats=False):
        W, H = view.width, view.height
        out = np.zeros((H * W, 4), np.float32)
        feats = np.zeros((H * W, NFEAT), np.float32) if want_feats else None
        active = np.zeros(H * W, np.int32)
        steps = np.zeros(H * W, np.int32)
        pos = np.zeros((H * W, 3), np.float32)
        lib().orc_render_predict(*self._vol(), ctypes.byref(view), _p(out), _p(feats), _p(active, _i),
                                 _p(steps, _i), _p(pos))
        return dict(out=out, feats=feats, active=active, steps=steps, pos=pos)

    def render_predict_rows(self, view, p0, n):
        out = np.zeros((n, 4), np.float32)
        lib().orc_render_predict_rows(*self._vol(), ctypes.byref(view), int(p0), int(n), _p(out))
        return out

    def render_predict_list(self, view, pix, want_feats=False):
        """NNPredict<MRPNN> on the listed pixels: (n, 4) outputs, and with want_feats also their
        (n, 577) features."""
        pix = np.ascontiguousarray(pix, np.int32)
        out = np.zeros((pix.size, 4), np.float32)
        feats = np.zeros((pix.size, NFEAT), np.float32) if want_feats else None
        lib().orc_render_predict_list(*self._vol(), ctypes.byref(view), _p(pix, _i), int(pix.size), _p(out),
                                      _p(feats))
        return (out, feats) if want_feats else out

    @classmethod
    def from_arrays(cls, grid, N, max_density, mips, tr, lut):
        """Scene over externally built pyramids (no CPU pyramid build)."""
        sc = cls.__new__(cls)
        sc.grid = np.ascontiguousarray(grid, np.float32)
        sc.N = N
        sc.max_density = float(max_density)
        sc.mips = np.ascontiguousarray(mips, np.float32)
        sc.tr = np.ascontiguousarray(tr, np.float32)
        sc.lut = np.ascontiguousarray(lut, np.float32)
        return sc

    def tile_gather(self, view, off_x=0, off_y=0):
        """GatherData (core/volume.cu:630-708): (H*W, 773) rows and the active flags."""
        W, H = view.width, view.height
        out = np.zeros((H * W, TILE_ROW), np.float32)
        active = np.zeros(H * W, np.int32)
        lib().orc_tile_gather(*self._vol(), ctypes.byref(view), int(off_x), int(off_y), _p(out), _p(active, _i))
        return out, active

    def gather_state(self, active, pos, X, Y, Z, light, g, alpha):
        out = np.zeros(NFEAT, np.float32)
        a = [np.asarray(v, np.float32) for v in (pos, X, Y, Z, light)]
        lib().orc_gather_state(*self._vol(), int(active), *[_p(x) for x in a], float(g), float(alpha), _p(out))
        return out

    def pt_render(self, view, spp, max_bounce, nee_baked=0, tr_scale=1.0):
        W, H = view.width, view.height
        mean = np.zeros((H * W, 3), np.float64)
        m2 = np.zeros((H * W, 3), np.float64)
        taps = ctypes.c_int64()
        lib().orc_pt_render(_p(self.grid), self.N, self.max_density, _p(self.tr), ctypes.byref(view), spp,
                            max_bounce, nee_baked, tr_scale, _p(mean, _d), _p(m2, _d), ctypes.byref(taps))
        return mean, m2, taps.value


def _pt_list(scene, view, pix, spp, max_bounce, nee_baked=0, tr_scale=1.0):
    pix = np.ascontiguousarray(pix, np.int32)
    mean = np.zeros((pix.size, 3), np.float64)
    m2 = np.zeros((pix.size, 3), np.float64)
    lib().orc_pt_render_list(_p(scene.grid), scene.N, scene.max_density, _p(scene.tr), ctypes.byref(view),
                             _p(pix, _i), int(pix.size), spp, max_bounce, nee_baked, tr_scale, _p(mean, _d),
                             _p(m2, _d))
    return mean, m2


Scene.pt_render_list = _pt_list


def _pt_list_cached(scene, view, pix, spp, max_bounce, ctr, ctr_res, tr_scale=1.0):
    """The path tracer on listed pixels with NEE from a cached transmittance bake (ctr_res^3)."""
    pix = np.ascontiguousarray(pix, np.int32)
    ctr = np.ascontiguousarray(ctr, np.float32)
    mean = np.zeros((pix.size, 3), np.float64)
    m2 = np.zeros((pix.size, 3), np.float64)
    lib().orc_pt_render_list_cached(_p(scene.grid), scene.N, scene.max_density, _p(ctr), int(ctr_res),
                                    ctypes.byref(view), _p(pix, _i), int(pix.size), spp, max_bounce, tr_scale,
                                    _p(mean, _d), _p(m2, _d))
    return mean, m2


def _bake_cached_tr(scene, light, alpha, tr_scale, seed, res, texels):
    """Listed texels (flat indices) of the Fill_BakeTR grid of resolution res."""
    t = np.ascontiguousarray(texels, np.int64)
    out = np.zeros(t.size, np.float32)
    L = np.asarray(light, np.float32)
    lib().orc_bake_cached_tr(_p(scene.grid), scene.N, scene.max_density, _p(L), float(alpha), float(tr_scale),
                             int(seed), int(res), t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(t.size),
                             _p(out))
    return out


def _point_radiance(scene, ori, dirs, lights, alpha, g, scatter, light_color, multi_scatter, spp, seed, tr_scale=1.0):
    """GetSamples: per task (mean, m2) over spp point-radiance paths, float64 [n, 3] each."""
    f = lambda a, k: np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, k))
    ori, dirs, lights = f(ori, 3), f(dirs, 3), f(lights, 3)
    alpha, g, scatter = f(alpha, 1), f(g, 1), f(scatter, 1)
    n = ori.shape[0]
    lc = np.asarray(light_color, np.float32)
    mean = np.zeros((n, 3), np.float64)
    m2 = np.zeros((n, 3), np.float64)
    lib().orc_point_radiance(_p(scene.grid), scene.N, scene.max_density, n, _p(ori), _p(dirs), _p(lights), _p(alpha),
                             _p(g), _p(scatter), _p(lc), int(multi_scatter), int(spp), int(seed), float(tr_scale),
                             _p(mean, _d), _p(m2, _d))
    return mean, m2


def _tr_samples(scene, ori, dirs, light, alpha, g, spp, seed, tr_scale=1.0):
    """ShadowTerm_TR(s): per task the mean over spp of Tr(ori -> light) * HG(dot(dir, light), g)."""
    ori = np.ascontiguousarray(np.asarray(ori, np.float32).reshape(-1, 3))
    dirs = np.ascontiguousarray(np.asarray(dirs, np.float32).reshape(-1, 3))
    out = np.zeros(ori.shape[0], np.float64)
    L = np.asarray(light, np.float32)
    lib().orc_tr_samples(_p(scene.grid), scene.N, scene.max_density, ori.shape[0], _p(ori), _p(dirs), _p(L),
                         float(alpha), float(g), int(spp), int(seed), float(tr_scale), _p(out, _d))
    return out


Scene.pt_render_list_cached = _pt_list_cached
Scene.bake_cached_tr = _bake_cached_tr
Scene.point_radiance = _point_radiance
Scene.tr_samples = _tr_samples


def sample_t(grid, N, max_density, ori, d, alpha):
    """GetSampleT (core/render.cuh:464-486): (t at survival 0.75, t at 0.25, first t, opacity)."""
    grid = np.ascontiguousarray(grid, np.float32)
    ts = np.zeros(4, np.float32)
    o, dd = np.asarray(ori, np.float32), np.asarray(d, np.float32)
    lib().orc_sample_t(_p(grid), N, max_density, _p(o), _p(dd), float(alpha), _p(ts))
    return ts


def march(grid, N, max_density, ori, d, alpha):
    grid = np.ascontiguousarray(grid, np.float32)
    pos = np.zeros(3, np.float32)
    act, taps = ctypes.c_int32(), ctypes.c_int32()
    o, dd = np.asarray(ori, np.float32), np.asarray(d, np.float32)
    lib().orc_march(_p(grid), N, max_density, _p(o), _p(dd), float(alpha), _p(pos), ctypes.byref(act),
                    ctypes.byref(taps))
    return bool(act.value), pos, taps.value


def basis(d, seed, frame, pixel, view_id):
    out = np.zeros(9, np.float32)
    dd = np.asarray(d, np.float32)
    lib().orc_basis(_p(dd), seed, frame, pixel, view_id, _p(out))
    return out.reshape(3, 3)


def tr_stochastic(grid, N, max_density, pos, d, dis, alpha, n, seed=7, tr_scale=1.0):
    grid = np.ascontiguousarray(grid, np.float32)
    out = np.zeros(n, np.float32)
    p, dd = np.asarray(pos, np.float32), np.asarray(d, np.float32)
    lib().orc_tr_stochastic(_p(grid), N, max_density, _p(p), _p(dd), float(dis), float(alpha), float(tr_scale),
                            seed, n, _p(out))
    return out


def orbit_view(width, height, dist=2.0, azimuth=0.0, elevation=0.0, fov_deg=40.0, light=(0.866, 0.5, 0.0),
               alpha=64.0, g=0.857, albedo=0.99, seed=1, frame=0, view_id=0, jitter=0.5, volume_index=0):
    """Camera on a sphere looking at the origin, basis as Camera::SetPosition (camera.cu:176-192)."""
    ca, sa, ce, se = np.cos(azimuth), np.sin(azimuth), np.cos(elevation), np.sin(elevation)
    ori = np.array([dist * ce * ca, dist * se, dist * ce * sa], np.float32)
    fwd = -ori / np.linalg.norm(ori)
    right = np.cross(np.array([0, 1, 0], np.float32), fwd)
    right = right / np.linalg.norm(right)
    up = np.cross(fwd, right)
    v = View()
    v.ori[:] = ori.tolist()
    v.up[:] = up.astype(np.float32).tolist()
    v.right[:] = right.astype(np.float32).tolist()
    L = np.asarray(light, np.float64)
    v.light_dir[:] = (L / np.linalg.norm(L)).astype(np.float32).tolist()
    v.light_color[:] = [1.0, 1.0, 1.0]
    v.scatter_rate[:] = [albedo] * 3
    v.alpha, v.g = alpha, g
    v.tan_half_fov = float(np.tan(np.radians(fov_deg) / 2))
    v.jitter = jitter
    v.width, v.height = width, height
    v.seed, v.frame, v.view_id, v.volume_index = seed, frame, view_id, volume_index
    return v


# ---------------------------------------------------------------------------------------
# Interactive progressive path (core/volume.cu:57-372, core/render.cuh:1144-1236).
# Images are (height*width, 3) float32 targets, (height*width, 33) float32 histograms
# (30 bins, total, x2, x) and (height*width, 4) float32 frame radiance; all in place.
# ---------------------------------------------------------------------------------------
HISTO_FLOATS = 33


def _c(a, dt):
    assert a.dtype == dt and a.flags["C_CONTIGUOUS"], "contiguous %s array required" % dt
    return a


def pt_accumulate(target, histo, rad4, W, H, checker, flip, fnum):
    _c(target, np.float32), _c(histo, np.float32), _c(rad4, np.float32)
    lib().orc_pt_accumulate(_p(target), histo.ctypes.data, _p(rad4), W, H, int(checker), int(flip), int(fnum))


def predict_commit(target, rad4, W, H, checker, flip):
    _c(target, np.float32), _c(rad4, np.float32)
    lib().orc_predict_commit(_p(target), _p(rad4), W, H, int(checker), int(flip))


def denoise(target, histo, W, H, frame_num, flip, denoise_on, exposure, tone):
    _c(target, np.float32), _c(histo, np.float32)
    rgba = np.zeros(W * H, np.uint32)
    lib().orc_denoise(_p(target), histo.ctypes.data, _p(rgba, _u), W, H, int(frame_num), int(flip), int(denoise_on),
                      float(exposure), int(tone))
    return rgba


def reprojection_denoise(target, histo, exposure, tone):
    _c(target, np.float32), _c(histo, np.float32)
    n = target.shape[0]
    rgba = np.zeros(n, np.uint32)
    lib().orc_reprojection_denoise(_p(target), histo.ctypes.data, _p(rgba, _u), n, float(exposure), int(tone))
    return rgba


def tonemap(rgb, exposure, tone):
    """(packed 0xAARRGGBB, the tone-mapped float rgb before saturation)."""
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    n = rgb.shape[0]
    rgba = np.zeros(n, np.uint32)
    mapped = np.zeros((n, 3), np.float32)
    lib().orc_tonemap(_p(rgb), n, float(exposure), int(tone), _p(rgba, _u), _p(mapped))
    return rgba, mapped


def histo_compare(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return float(lib().orc_histo_compare(x.ctypes.data, y.ctypes.data))


# ---------------------------------------------------------------------------------------
# Environment map (SetHDRI / SetEnvExp, SkyBox core/render.cuh:24-27): applies to every later
# oracle call until removed with set_env(None).
# ---------------------------------------------------------------------------------------
_env_keep = None


def set_env(rgba, env_exp=1.0):
    global _env_keep
    if rgba is None:
        _env_keep = None
        lib().orc_set_env(None, 0, 0, float(env_exp))
        return
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    _env_keep = rgba
    lib().orc_set_env(_p(rgba), int(w), int(h), float(env_exp))


def sky_box(d):
    out = np.zeros(3, np.float32)
    dd = np.ascontiguousarray(d, np.float32)
    lib().orc_sky_box(_p(dd), _p(out))
    return out


# ---------------------------------------------------------------------------------------
# RPNN comparison predictor (core/radiancePredict.cu:567-722)
# ---------------------------------------------------------------------------------------
_rpnn_loaded = False


def _rpnn():
    global _rpnn_loaded
    L = lib()
    if not _rpnn_loaded:
        w = np.fromfile(RPNN_WEIGHTS, dtype="<f4")
        if L.orc_rpnn_load(w.ctypes.data_as(_f), int(w.size)) != 0:
            raise RuntimeError("oracle: RPNN weight asset does not match the network layout")
        _rpnn_loaded = True
    return L


def rpnn_predict_f32(rows):
    rows = np.ascontiguousarray(rows, np.float32).reshape(-1, RPNN_ROW)
    out = np.zeros(rows.shape[0], np.float32)
    _rpnn().orc_rpnn_predict_f32(_p(rows), rows.shape[0], _p(out))
    return out


def rpnn_predict_f64(rows):
    rows = np.ascontiguousarray(rows, np.float32).reshape(-1, RPNN_ROW)
    out = np.zeros(rows.shape[0], np.float64)
    _rpnn().orc_rpnn_predict_f64(_p(rows), rows.shape[0], _p(out, _d))
    return out


def rpnn_descriptor(scene, pos, X, LX, alpha):
    desc = np.zeros(2250, np.float32)
    gamma = np.zeros(1, np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (pos, X, LX)]
    _rpnn().orc_rpnn_descriptor(_p(scene.grid), scene.N, scene.max_density, _p(scene.mips), *[_p(x) for x in a],
                                float(alpha), _p(desc), _p(gamma))
    return desc, float(gamma[0])


def render_rpnn(scene, view):
    W, H = view.width, view.height
    rows = np.zeros((H * W, RPNN_ROW), np.float32)
    active = np.zeros(H * W, np.int32)
    out = np.zeros((H * W, 4), np.float32)
    _rpnn().orc_render_rpnn(_p(scene.grid), scene.N, scene.max_density, _p(scene.mips), ctypes.byref(view),
                            _p(rows), _p(active, _i), _p(out))
    return dict(rows=rows, active=active, out=out)
