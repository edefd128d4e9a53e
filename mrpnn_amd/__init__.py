"""Python binding of libmrpnn_hip.so (the C ABI in include/mrp_c_api.h).

The library is the product: gfx950 HIP kernels behind a C ABI.  This module is a thin
ctypes layer for tests and benchmarks; device buffers are torch tensors (plumbing only).
There is no CPU fallback: if the library is missing, loading fails loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrpnn_hip.so")

FP32, BF16, FP8, BF16_1T = 0, 1, 2, 3
PRECISIONS = {"fp32": FP32, "bf16": BF16, "fp8": FP8, "bf16_1t": BF16_1T}
OPT_MFMA_WAVES_PER_CU, OPT_EXACT_GATHER, OPT_TR_SLOTS, OPT_GRAPH, OPT_LDS_WEIGHTS, OPT_RPNN_CHUNK = 1, 2, 3, 4, 5, 6
OPT_TILE_W = 7
NFEAT = 577
FP8_LAYERS = 60  # MFMA layers with an FP8 activation exponent (MRP_FP8_LAYERS)
TILE_ROW = 773  # tile feature gather row (mrp_gather_tiles)
MIP_TOTAL = sum((256 >> m) ** 3 for m in range(9))
TR_TOTAL = sum((128 >> m) ** 3 for m in range(8))


class MrpError(RuntimeError):
    pass


class View(ctypes.Structure):
    """mrp_view (include/mrp_c_api.h)."""
    _fields_ = [("ori", ctypes.c_float * 3), ("up", ctypes.c_float * 3), ("right", ctypes.c_float * 3),
                ("light_dir", ctypes.c_float * 3), ("light_color", ctypes.c_float * 3),
                ("scatter_rate", ctypes.c_float * 3), ("alpha", ctypes.c_float), ("g", ctypes.c_float),
                ("tan_half_fov", ctypes.c_float), ("jitter", ctypes.c_float), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("seed", ctypes.c_uint32), ("frame", ctypes.c_uint32),
                ("view_id", ctypes.c_uint32), ("volume_index", ctypes.c_int32)]

    def copy(self):
        v = View()
        ctypes.pointer(v)[0] = self
        return v


class Stats(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_int64), ("active_pixels", ctypes.c_int64), ("march_taps", ctypes.c_int64),
                ("ms_march", ctypes.c_float), ("ms_predict", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("timed_renders", ctypes.c_int64), ("ms_march_sum", ctypes.c_double),
                ("ms_predict_sum", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("plain", ctypes.c_int32), ("max_batch_views", ctypes.c_int32),
                ("max_batch_pixels", ctypes.c_int64), ("max_inflight", ctypes.c_int32), ("g", ctypes.c_float),
                ("hbm_limit_bytes", ctypes.c_int64), ("hbm_reserve_bytes", ctypes.c_int64)]


class Request(ctypes.Structure):
    _fields_ = [("volume_id", ctypes.c_int32), ("view", View), ("dev_out4", ctypes.c_void_p),
                ("user_tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    _fields_ = [("user_tag", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_launch", ctypes.c_double),
                ("t_done", ctypes.c_double), ("status", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "pixels", "batches", "batched_views",
                                              "admissions", "evictions", "resident_volumes", "resident_bytes",
                                              "pending")]


class LoadSpec(ctypes.Structure):
    _fields_ = [("rate_rps", ctypes.c_double), ("duration_s", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n_sizes", ctypes.c_int32), ("sizes", ctypes.c_int32 * 16), ("buffers_per_size", ctypes.c_int32)]


class LoadReport(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("completed", ctypes.c_int64), ("pixels", ctypes.c_int64),
                ("wall_s", ctypes.c_double), ("offered_s", ctypes.c_double), ("mpix_s", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double)]


_lib = None
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MrpError("libmrpnn_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.mrp_last_error.restype = ctypes.c_char_p
        L.mrp_version.restype = ctypes.c_char_p
        L.mrp_volume_bytes.restype = ctypes.c_int64
        L.mrp_volume_bytes.argtypes = [ctypes.c_int]
        L.mrp_ctx_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(_vp)]
        L.mrp_ctx_destroy.argtypes = [_vp]
        L.mrp_ctx_set_stream.argtypes = [_vp, _vp]
        L.mrp_ctx_set_timing.argtypes = [_vp, ctypes.c_int]
        L.mrp_ctx_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
        L.mrp_ctx_update_hg_lut.argtypes = [_vp, ctypes.c_float, ctypes.c_uint32]
        L.mrp_ctx_download_hg_lut.argtypes = [_vp, _fp]
        L.mrp_ctx_upload_hg_lut.argtypes = [_vp, _fp, ctypes.c_float]
        L.mrp_ctx_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.mrp_server_create.argtypes = [_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
        L.mrp_server_destroy.argtypes = [_vp]
        L.mrp_server_add_volume.argtypes = [_vp, ctypes.c_int, _fp, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
        L.mrp_server_add_fbm_volume.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                                                ctypes.POINTER(ctypes.c_int)]
        L.mrp_server_submit.argtypes = [_vp, ctypes.POINTER(Request)]
        L.mrp_server_step.argtypes = [_vp]
        L.mrp_server_drain.argtypes = [_vp]
        L.mrp_server_poll.argtypes = [_vp, ctypes.POINTER(Completion), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.mrp_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.mrp_server_is_resident.argtypes = [_vp, ctypes.c_int]
        L.mrp_server_run_poisson.argtypes = [_vp, ctypes.POINTER(LoadSpec), ctypes.POINTER(LoadReport)]
        L.mrp_autotune_batch.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(View), ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_float)]
        L.mrp_ctx_get_option.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.mrp_volume_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.mrp_volume_destroy.argtypes = [_vp]
        L.mrp_volume_set_data.argtypes = [_vp, _fp, ctypes.c_float]
        L.mrp_volume_set_data_device.argtypes = [_vp, _vp, ctypes.c_float]
        L.mrp_volume_fill_fbm.argtypes = [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint32]
        L.mrp_volume_get_data.argtypes = [_vp, _fp]
        L.mrp_volume_update.argtypes = [_vp]
        L.mrp_volume_update_tr.argtypes = [_vp, _fp, ctypes.c_float]
        L.mrp_volume_download_mips.argtypes = [_vp, _fp]
        L.mrp_volume_download_tr.argtypes = [_vp, _fp]
        L.mrp_volume_bake_cached_tr.argtypes = [_vp, _fp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_int]
        L.mrp_volume_download_cached_tr.argtypes = [_vp, _fp]
        L.mrp_point_radiance.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _fp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_uint32, ctypes.c_float, _vp, _vp]
        L.mrp_tr_samples.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp, _fp, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_uint32, ctypes.c_float, _vp]
        L.mrp_render_plain.argtypes = [_vp, _vp, ctypes.POINTER(View), ctypes.c_int, _fp]
        L.mrp_render_plain_device.argtypes = [_vp, _vp, ctypes.POINTER(View), _vp]
        L.mrp_prepare_views.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(View), ctypes.c_int]
        L.mrp_render_batch.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(View), ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(_vp)]
        L.mrp_render_batch_noalloc.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(View),
                                               ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.mrp_gather_features.argtypes = [_vp, _vp, ctypes.POINTER(View), _vp, _vp, _vp, _vp]
        L.mrp_gather_tiles.argtypes = [_vp, _vp, ctypes.POINTER(View), ctypes.c_int, ctypes.c_int, _vp]
        L.mrp_gather_features_fast.argtypes = [_vp, _vp, ctypes.POINTER(View), _vp]
        L.mrp_predict_features.argtypes = [_vp, _vp, ctypes.c_int, _fp, ctypes.c_float, ctypes.c_int, _vp]
        L.mrp_fp8_calibrate.argtypes = [_vp, _vp, ctypes.c_int, _fp, ctypes.c_float, ctypes.c_int, _fp]
        L.mrp_fp8_get_exponents.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.mrp_fp8_set_exponents.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.mrp_pt_render.argtypes = [_vp, _vp, ctypes.POINTER(View), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_float, _vp, _vp]
        L.mrp_volume_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_fp), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_float)]
        L.mrp_host_free.argtypes = [_vp]
        L.mrp_host_free.restype = None
        L.mrp_volume_load.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float)]
        L.mrp_hdr_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_fp), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]
        L.mrp_ctx_set_env_map.argtypes = [_vp, _fp, ctypes.c_int, ctypes.c_int]
        L.mrp_ctx_load_hdri.argtypes = [_vp, ctypes.c_char_p]
        L.mrp_ctx_set_env_exp.argtypes = [_vp, ctypes.c_float]
        L.mrp_env_light.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp, _fp, ctypes.c_int,
                                    ctypes.c_float, _fp, _fp]
        L.mrp_ctx_env_light.argtypes = [_vp, _fp, _fp, ctypes.c_int, ctypes.c_float, _fp, _fp]
        L.mrp_ctx_load_rpnn.argtypes = [_vp, ctypes.c_char_p]
        L.mrp_render_rpnn.argtypes = [_vp, _vp, ctypes.POINTER(View), ctypes.c_int, _vp]
        L.mrp_rpnn_features.argtypes = [_vp, _vp, ctypes.POINTER(View), _vp, _vp]
        L.mrp_rpnn_predict_rows.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
        L.mrp_interactive_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.mrp_interactive_destroy.argtypes = [_vp]
        L.mrp_interactive_destroy.restype = None
        L.mrp_interactive_set_checkerboard.argtypes = [_vp, ctypes.c_int]
        L.mrp_interactive_set_exposure.argtypes = [_vp, ctypes.c_float]
        L.mrp_interactive_radiance.argtypes = [_vp, _vp]
        L.mrp_render_interactive.argtypes = [_vp, _vp, ctypes.POINTER(View), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise MrpError("mrp error %d: %s" % (rc, lib().mrp_last_error().decode()))


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    def __init__(self, weights_path=None):
        import weakref
        self.h = _vp()
        self._children = weakref.WeakSet()  # volumes and servers: closed before the context
        check(lib().mrp_ctx_create(weights_path.encode() if weights_path else None, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            for ch in list(self._children):
                if isinstance(ch, Server):
                    ch.close()
            for ch in list(self._children):
                ch.close()
            lib().mrp_ctx_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        check(lib().mrp_ctx_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def set_timing(self, on):
        check(lib().mrp_ctx_set_timing(self.h, 1 if on else 0))

    def stats(self):
        s = Stats()
        check(lib().mrp_ctx_get_stats(self.h, ctypes.byref(s)))
        return s

    def set_option(self, option, value):
        check(lib().mrp_ctx_set_option(self.h, int(option), int(value)))

    def load_rpnn(self, path=None):
        check(lib().mrp_ctx_load_rpnn(self.h, os.fsencode(path) if path else None))

    def rpnn_predict_rows(self, rows, precision=FP32):
        """The RPNN network on (n, 2251) float32 device rows: (n,) outputs."""
        import torch
        n = rows.shape[0]
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        check(lib().mrp_rpnn_predict_rows(self.h, _ptr(rows.contiguous()), int(n), int(precision), _ptr(y)))
        return y

    def set_env_map(self, rgba):
        """Environment map (height, width, 4) float32, or None to remove it (SetHDRI)."""
        if rgba is None:
            check(lib().mrp_ctx_set_env_map(self.h, None, 0, 0))
            return
        img = np.ascontiguousarray(rgba, np.float32)
        check(lib().mrp_ctx_set_env_map(self.h, img.ctypes.data_as(_fp), int(img.shape[1]), int(img.shape[0])))

    def load_hdri(self, path):
        check(lib().mrp_ctx_load_hdri(self.h, os.fsencode(path)))

    def set_env_exp(self, exposure):
        check(lib().mrp_ctx_set_env_exp(self.h, float(exposure)))

    def get_option(self, option):
        v = ctypes.c_int()
        check(lib().mrp_ctx_get_option(self.h, int(option), ctypes.byref(v)))
        return v.value

    def set_mfma_waves(self, w):
        self.set_option(OPT_MFMA_WAVES_PER_CU, w)

    def set_exact_gather(self, on):
        self.set_option(OPT_EXACT_GATHER, 1 if on else 0)

    def update_hg_lut(self, g, seed=1):
        check(lib().mrp_ctx_update_hg_lut(self.h, float(g), int(seed)))

    def download_hg_lut(self):
        out = np.zeros(512 * 512, np.float32)
        check(lib().mrp_ctx_download_hg_lut(self.h, out.ctypes.data_as(_fp)))
        return out

    def upload_hg_lut(self, lut, g):
        lut = np.ascontiguousarray(lut, np.float32)
        check(lib().mrp_ctx_upload_hg_lut(self.h, lut.ctypes.data_as(_fp), float(g)))

    def predict_features(self, feats, sr, g, precision=FP32):
        """feats: torch float32 [n, 577] on the GPU -> torch [n, 3]."""
        import torch
        n = feats.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=feats.device)
        srv = (ctypes.c_float * 3)(*[float(x) for x in sr])
        check(lib().mrp_predict_features(self.h, _ptr(feats), n, srv, float(g), int(precision), _ptr(out)))
        return out

    def fp8_calibrate(self, feats, sr, g, reset=True):
        """mrp_fp8_calibrate over feature rows (torch float32 [n, 577] on the GPU, or None to
        only re-derive the exponents); returns the per-layer input maxima (numpy, 60)."""
        srv = (ctypes.c_float * 3)(*[float(x) for x in sr])
        mx = (ctypes.c_float * FP8_LAYERS)()
        n = 0 if feats is None else feats.shape[0]
        check(lib().mrp_fp8_calibrate(self.h, None if feats is None else _ptr(feats), n, srv, float(g), int(reset), mx))
        return np.frombuffer(mx, np.float32).copy()

    def fp8_exponents(self):
        e = (ctypes.c_int * FP8_LAYERS)()
        check(lib().mrp_fp8_get_exponents(self.h, e, FP8_LAYERS))
        return np.frombuffer(e, np.int32).copy()

    def set_fp8_exponents(self, exps):
        e = (ctypes.c_int * FP8_LAYERS)(*[int(x) for x in exps])
        check(lib().mrp_fp8_set_exponents(self.h, e, len(exps)))

    def render_batch(self, volumes, views, outs, precision=BF16):
        vh = (_vp * len(volumes))(*[v.h for v in volumes])
        va = (View * len(views))(*views)
        oa = (_vp * len(outs))(*[o.data_ptr() for o in outs])
        check(lib().mrp_render_batch(self.h, vh, len(volumes), va, len(views), int(precision), oa))

    def render_batch_noalloc(self, volumes, views, outs, precision=BF16):
        """mrp_render_batch_noalloc: the same batch with no device allocation (static scratch)."""
        vh = (_vp * len(volumes))(*[v.h for v in volumes])
        va = (View * len(views))(*views)
        oa = (_vp * len(outs))(*[o.data_ptr() for o in outs])
        check(lib().mrp_render_batch_noalloc(self.h, vh, len(volumes), va, len(views), int(precision), oa))

    def autotune(self, volumes, views, outs, precision=BF16):
        """Times the prepared batch per candidate waves-per-CU; keeps and returns the best
        (waves_per_cu, ms per batch)."""
        vh = (_vp * len(volumes))(*[v.h for v in volumes])
        va = (View * len(views))(*views)
        oa = (_vp * len(outs))(*[o.data_ptr() for o in outs])
        w, ms = ctypes.c_int(), ctypes.c_float()
        check(lib().mrp_autotune_batch(self.h, vh, len(volumes), va, len(views), int(precision), oa,
                                       ctypes.byref(w), ctypes.byref(ms)))
        return w.value, ms.value

    def prepare_views(self, volumes, views):
        vh = (_vp * len(volumes))(*[v.h for v in volumes])
        va = (View * len(views))(*views)
        check(lib().mrp_prepare_views(self.h, vh, len(volumes), va, len(views)))


class Volume:
    def __init__(self, ctx, N, _handle=None):
        self.ctx = ctx
        self.N = N
        self.h = _vp()
        if _handle is not None:
            self.h = _handle
        else:
            check(lib().mrp_volume_create(ctx.h, int(N), ctypes.byref(self.h)))
        ctx._children.add(self)

    @classmethod
    def load(cls, ctx, path):
        """VolumeRender::LoadData (core/volume.cu:970-1149) through mrp_volume_load: a named
        dataset (CLOUD0/1, MODEL0/1) from <path>.bin, or a .txt / .vox / .bin grid, uploaded and
        its pyramid built."""
        h = _vp()
        md = ctypes.c_float()
        check(lib().mrp_volume_load(ctx.h, os.fsencode(path), ctypes.byref(h), ctypes.byref(md)))
        vol = cls(ctx, 0, _handle=h)
        vol.max_density = md.value
        vol.path = str(path)
        return vol

    def close(self):
        if self.h:
            lib().mrp_volume_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, grid, max_density):
        grid = np.ascontiguousarray(grid, np.float32)
        assert grid.size == self.N ** 3
        check(lib().mrp_volume_set_data(self.h, grid.ctypes.data_as(_fp), float(max_density)))

    def fill_fbm(self, octaves, base_freq, seed):
        check(lib().mrp_volume_fill_fbm(self.h, int(octaves), float(base_freq), int(seed)))

    def get_data(self):
        out = np.zeros(self.N ** 3, np.float32)
        check(lib().mrp_volume_get_data(self.h, out.ctypes.data_as(_fp)))
        return out

    def update(self):
        check(lib().mrp_volume_update(self.h))

    def update_tr(self, light, alpha):
        L = (ctypes.c_float * 3)(*[float(x) for x in light])
        check(lib().mrp_volume_update_tr(self.h, L, float(alpha)))

    def download_mips(self):
        out = np.zeros(MIP_TOTAL, np.float32)
        check(lib().mrp_volume_download_mips(self.h, out.ctypes.data_as(_fp)))
        return out

    def download_tr(self):
        out = np.zeros(TR_TOTAL, np.float32)
        check(lib().mrp_volume_download_tr(self.h, out.ctypes.data_as(_fp)))
        return out

    def render_plain(self, view, n_frames=1):
        out = np.zeros((view.height * view.width, 4), np.float32)
        check(lib().mrp_render_plain(self.ctx.h, self.h, ctypes.byref(view), int(n_frames), out.ctypes.data_as(_fp)))
        return out

    def render_rpnn(self, view, precision=FP32, out=None):
        """RPNN frame (mrp_render_rpnn): (height*width, 4) float32 on the GPU."""
        import torch
        if out is None:
            out = torch.zeros((view.width * view.height, 4), dtype=torch.float32, device="cuda")
        check(lib().mrp_render_rpnn(self.ctx.h, self.h, ctypes.byref(view), int(precision), _ptr(out)))
        return out

    def rpnn_features(self, view):
        """(rows (height*width, 2251), out4 with the sky pixels) as the RPNN path computes them."""
        import torch
        n = view.width * view.height
        rows = torch.empty((n, 2251), dtype=torch.float32, device="cuda")
        out = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        check(lib().mrp_rpnn_features(self.ctx.h, self.h, ctypes.byref(view), _ptr(rows), _ptr(out)))
        return rows, out

    def render_plain_device(self, view, out):
        check(lib().mrp_render_plain_device(self.ctx.h, self.h, ctypes.byref(view), _ptr(out)))

    def gather_features(self, view):
        import torch
        n = view.width * view.height
        dev = torch.device("cuda")
        feats = torch.empty((n, NFEAT), dtype=torch.float32, device=dev)
        active = torch.empty(n, dtype=torch.int32, device=dev)
        pos = torch.empty((n, 3), dtype=torch.float32, device=dev)
        taps = torch.empty(n, dtype=torch.int32, device=dev)
        check(lib().mrp_gather_features(self.ctx.h, self.h, ctypes.byref(view), _ptr(feats), _ptr(active), _ptr(pos),
                                        _ptr(taps)))
        return feats, active, pos, taps

    def gather_features_fast(self, view):
        """mrp_gather_features_fast: (height*width, 577) float32 features of the fused kernels'
        default gather on the GPU (zero rows where no first-scattering point exists)."""
        import torch
        feats = torch.zeros((view.width * view.height, NFEAT), dtype=torch.float32, device="cuda")
        check(lib().mrp_gather_features_fast(self.ctx.h, self.h, ctypes.byref(view), _ptr(feats)))
        return feats

    def gather_tiles(self, view, off_x=0, off_y=0, out=None):
        """Tile feature gather (mrp_gather_tiles): (height*width, 773) float32 rows on the GPU."""
        import torch
        n = view.width * view.height
        if out is None:
            out = torch.empty((n, TILE_ROW), dtype=torch.float32, device="cuda")
        if out.numel() < n * TILE_ROW or not out.is_contiguous() or out.dtype != torch.float32:
            raise MrpError("gather_tiles: out must be a contiguous float32 tensor of %d x %d" % (n, TILE_ROW))
        check(lib().mrp_gather_tiles(self.ctx.h, self.h, ctypes.byref(view), int(off_x), int(off_y), _ptr(out)))
        return out

    def bake_cached_tr(self, light, alpha, tr_scale=1.0, seed=1, res=0):
        """mrp_volume_bake_cached_tr: the path tracer's cached transmittance (res^3, 0 = 768)."""
        L = (ctypes.c_float * 3)(*[float(x) for x in light])
        check(lib().mrp_volume_bake_cached_tr(self.h, L, float(alpha), float(tr_scale), int(seed), int(res)))

    def download_cached_tr(self, res):
        out = np.zeros(res ** 3, np.float32)
        check(lib().mrp_volume_download_cached_tr(self.h, out.ctypes.data_as(_fp)))
        return out

    def point_radiance(self, ori, dirs, lights, alpha, g, scatter, light_color=(1, 1, 1), multi_scatter=512,
                       spp=1024, seed=1, tr_scale=1.0):
        """mrp_point_radiance (GetSamples): per task the mean radiance (torch float32 [n, 3]) and
        its FP64 M2 (torch float64 [n, 3]) on the GPU.  Inputs: array-likes of n rows."""
        import torch
        dev = torch.device("cuda")
        t = lambda a, k: torch.as_tensor(np.asarray(a, np.float32).reshape(-1, k), device=dev).contiguous()
        ori, dirs, lights = t(ori, 3), t(dirs, 3), t(lights, 3)
        alpha, g, scatter = t(alpha, 1), t(g, 1), t(scatter, 1)
        n = ori.shape[0]
        out = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        m2 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        lc = (ctypes.c_float * 3)(*[float(x) for x in light_color])
        check(lib().mrp_point_radiance(self.ctx.h, self.h, n, _ptr(ori), _ptr(dirs), _ptr(lights), _ptr(alpha), _ptr(g),
                                       _ptr(scatter), lc, int(multi_scatter), int(spp), int(seed), float(tr_scale),
                                       _ptr(out), _ptr(m2)))
        return out, m2

    def tr_samples(self, ori, dirs, light, alpha, g, spp, seed=1, tr_scale=1.0):
        """mrp_tr_samples: per task the mean of Tr(ori -> light) * HG(dot(dir, light), g)."""
        import torch
        dev = torch.device("cuda")
        ori = torch.as_tensor(np.asarray(ori, np.float32).reshape(-1, 3), device=dev).contiguous()
        dirs = torch.as_tensor(np.asarray(dirs, np.float32).reshape(-1, 3), device=dev).contiguous()
        out = torch.zeros(ori.shape[0], dtype=torch.float32, device=dev)
        L = (ctypes.c_float * 3)(*[float(x) for x in light])
        check(lib().mrp_tr_samples(self.ctx.h, self.h, ori.shape[0], _ptr(ori), _ptr(dirs), L, float(alpha), float(g),
                                   int(spp), int(seed), float(tr_scale), _ptr(out)))
        return out

    def pt_render(self, view, spp, max_bounce, nee_baked=0, tr_scale=1.0, chunk=None, mean=None, m2=None, spp0=0):
        import torch
        n = view.width * view.height
        if mean is None:
            mean = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
            m2 = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        chunk = chunk or spp
        done = 0
        while done < spp:
            k = min(chunk, spp - done)
            check(lib().mrp_pt_render(self.ctx.h, self.h, ctypes.byref(view), spp0 + done, k, int(max_bounce),
                                      int(nee_baked), float(tr_scale), _ptr(mean), _ptr(m2)))
            done += k
        return mean, m2


def orbit_view(width, height, dist=2.0, azimuth=0.0, elevation=0.0, fov_deg=40.0, light=(0.866, 0.5, 0.0),
               alpha=64.0, g=0.857, albedo=0.99, seed=1, frame=0, view_id=0, jitter=0.5, volume_index=0):
    """Camera on a sphere looking at the origin; basis as Camera::SetPosition (camera.cu:176-192)."""
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


def camera_view(width, height, position, light_dir, alpha=1.0, g=0.857, fov_deg=90.0):
    """View of Camera(VolumeRender&, name, position) after SetPosition (core/camera.cu:166-192):
    looking at the origin, right = normalize(cross(y, forward)), up = cross(forward, right); the
    reference's camera has a fixed 90 degree field of view."""
    ori = np.asarray(position, np.float32)
    fwd = -ori / np.linalg.norm(ori)
    right = np.cross(np.array([0, 1, 0], np.float32), fwd)
    right = right / np.linalg.norm(right)
    up = np.cross(fwd, right)
    v = View()
    v.ori[:] = ori.tolist()
    v.up[:] = up.astype(np.float32).tolist()
    v.right[:] = right.astype(np.float32).tolist()
    L = np.asarray(light_dir, np.float64)
    v.light_dir[:] = (L / np.linalg.norm(L)).astype(np.float32).tolist()
    v.light_color[:] = [1.0, 1.0, 1.0]
    v.scatter_rate[:] = [1.0, 1.0, 1.0]
    v.alpha, v.g = float(alpha), float(g)
    v.tan_half_fov = float(np.tan(np.radians(fov_deg) / 2))
    v.jitter = 0.5
    v.width, v.height = int(width), int(height)
    v.seed, v.frame, v.view_id, v.volume_index = 1, 0, 0, 0
    return v


RT_PT, RT_RPNN, RT_MRPNN = 0, 1, 2  # VolumeRender::RenderType
TONE_NONE, TONE_GAMMA, TONE_ACES = 0, 1, 2
HISTO_FLOATS = 33  # mrp_histogram: 30 bins, total sample count, x2, x


class Interactive:
    """Interactive progressive renderer (mrp_render_interactive): VolumeRender::Render(float3*
    target, Histogram*, unsigned* target2, ...) (core/volume.cu:550-628).  The caller owns the
    device buffers: target (height*width, 3) float32, histo (height*width, 33) float32 and
    rgba (height*width,) int32 holding 0xAARRGGBB words."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, int(width), int(height)
        self.h = _vp()
        check(lib().mrp_interactive_create(ctx.h, self.width, self.height, ctypes.byref(self.h)))
        ctx._children.add(self)

    def close(self):
        if self.h:
            lib().mrp_interactive_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buffers(self):
        import torch
        n = self.width * self.height
        return (torch.zeros((n, 3), dtype=torch.float32, device="cuda"),
                torch.zeros((n, HISTO_FLOATS), dtype=torch.float32, device="cuda"),
                torch.zeros((n,), dtype=torch.int32, device="cuda"))

    def set_checkerboard(self, on):
        check(lib().mrp_interactive_set_checkerboard(self.h, 1 if on else 0))

    def set_exposure(self, exposure):
        check(lib().mrp_interactive_set_exposure(self.h, float(exposure)))

    def render(self, volume, view, target, histo, rgba, render_type=RT_MRPNN, precision=FP32, max_bounce=1,
               randseed=0, tone=TONE_ACES, denoise=False):
        n = self.width * self.height
        for t, k, name in ((target, 3 * n, "target"), (histo, HISTO_FLOATS * n, "histo"), (rgba, n, "rgba")):
            if t.numel() < k or not t.is_contiguous() or t.element_size() != 4:
                raise MrpError("render_interactive: %s must be a contiguous 4-byte tensor of %d elements" % (name, k))
        check(lib().mrp_render_interactive(self.h, volume.h, ctypes.byref(view), int(render_type), int(precision),
                                           int(max_bounce), int(randseed), int(tone), 1 if denoise else 0,
                                           _ptr(target), _ptr(histo), _ptr(rgba)))

    def radiance(self):
        """The last frame's radiance (height*width, 4): rgb after max(0), distance; w = -2 where the
        frame did not render (checkerboard)."""
        import torch
        out = torch.empty((self.width * self.height, 4), dtype=torch.float32, device="cuda")
        check(lib().mrp_interactive_radiance(self.h, _ptr(out)))
        return out


def read_volume(path):
    """Host-side grid of `path` as LoadData reads it: (float32 N^3 array (x slowest), N,
    max_density).  Writes <path>.bin after a .txt / .vox parse, as the reference."""
    p = _fp()
    n = ctypes.c_int()
    md = ctypes.c_float()
    check(lib().mrp_volume_read(os.fsencode(path), ctypes.byref(p), ctypes.byref(n), ctypes.byref(md)))
    try:
        grid = np.ctypeslib.as_array(p, shape=(n.value ** 3,)).copy()
    finally:
        lib().mrp_host_free(p)
    return grid, n.value, md.value


def read_hdr(path):
    """Radiance RGBE image as (height, width, 4) float32 (alpha 0), rows in file order."""
    p = _fp()
    w, h = ctypes.c_int(), ctypes.c_int()
    check(lib().mrp_hdr_read(os.fsencode(path), ctypes.byref(p), ctypes.byref(w), ctypes.byref(h)))
    try:
        img = np.ctypeslib.as_array(p, shape=(h.value * w.value * 4,)).copy()
    finally:
        lib().mrp_host_free(p)
    return img.reshape(h.value, w.value, 4)


def env_light(rgba, env_exp, light_dir, light_color, counter, u):
    """The per-frame choice between the sun and an environment direction (mrp_env_light):
    (chose_env, direction, colour)."""
    img = None if rgba is None else np.ascontiguousarray(rgba, np.float32)
    d = np.asarray(light_dir, np.float32)
    c = np.asarray(light_color, np.float32)
    od = np.zeros(3, np.float32)
    oc = np.zeros(3, np.float32)
    h, w = (img.shape[0], img.shape[1]) if img is not None else (0, 0)
    r = lib().mrp_env_light(img.ctypes.data_as(_fp) if img is not None else None, w, h, float(env_exp),
                            d.ctypes.data_as(_fp), c.ctypes.data_as(_fp), int(counter), float(u),
                            od.ctypes.data_as(_fp), oc.ctypes.data_as(_fp))
    return bool(r), od, oc


class MRPNNInputData:
    """The reference's training-input fetcher (mrpnn_input_data_wrapper.cpp:34-103): a volume,
    a camera at 1/8 of the image resolution and a (height/8, width/8, 773) device tensor that
    fetch_input refills with mrp_gather_tiles.  The reference takes a volume file path; here a
    resident `Volume` (swap_data replaces it).  Unlike the reference's square camera
    (resolution = width/8 for both axes), the tile grid is width/8 x height/8."""

    def __init__(self, width, height, volume):
        import torch
        if isinstance(volume, (str, bytes, os.PathLike)):
            raise MrpError("MRPNNInputData: pass a Volume, or use MRPNNInputData.from_path(width, height, ctx, path)")
        self.volume = volume
        self.w, self.h = int(width) // 8, int(height) // 8
        if self.w < 1 or self.h < 1:
            raise MrpError("MRPNNInputData: width and height must be at least 8")
        self.tensor = torch.empty((self.h, self.w, TILE_ROW), dtype=torch.float32, device="cuda")

    def fetch_input(self, camPos=(0.67085, -0.03808, -0.04856), lightDir=(0.34281, 0.70711, 0.61845), alpha=1.0,
                    g=0.857, pixelOffset=(0, 0)):
        v = camera_view(self.w, self.h, camPos, lightDir, alpha, g)
        self.volume.gather_tiles(v, int(pixelOffset[0]), int(pixelOffset[1]), out=self.tensor)
        return self.tensor

    @classmethod
    def from_path(cls, width, height, ctx, vol_path):
        """The reference's constructor MRPNNInputData(width, height, vol_path)
        (mrpnn_input_data_wrapper.cpp:43-67): the volume loaded as LoadData does."""
        return cls(width, height, Volume.load(ctx, vol_path))

    def swap_data(self, volume):
        """A resident Volume, or a path loaded as LoadData when it differs from the current
        one (mrpnn_input_data_wrapper.cpp:94-101)."""
        if isinstance(volume, (str, os.PathLike)):
            if getattr(self.volume, "path", None) == str(volume):
                return
            volume = Volume.load(self.volume.ctx, volume)
        self.volume = volume


class Server:
    """Serving loop over resident volumes (mrp_server_* in include/mrp_c_api.h)."""

    def __init__(self, ctx, precision=BF16, plain=False, max_batch_views=64, max_batch_pixels=16 << 20,
                 max_inflight=2, g=0.857, hbm_limit_bytes=0, hbm_reserve_bytes=0):
        self.ctx = ctx
        cfg = ServerConfig(int(precision), 1 if plain else 0, int(max_batch_views), int(max_batch_pixels),
                           int(max_inflight), float(g), int(hbm_limit_bytes), int(hbm_reserve_bytes))
        self.h = _vp()
        check(lib().mrp_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(self.h)))
        ctx._children.add(self)

    def close(self):
        if self.h:
            lib().mrp_server_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_volume(self, grid, N, max_density=1.0):
        grid = np.ascontiguousarray(grid, np.float32)
        vid = ctypes.c_int()
        check(lib().mrp_server_add_volume(self.h, int(N), grid.ctypes.data_as(_fp), float(max_density),
                                          ctypes.byref(vid)))
        return vid.value

    def add_fbm_volume(self, N, octaves=5, base_freq=4.0, seed=1):
        vid = ctypes.c_int()
        check(lib().mrp_server_add_fbm_volume(self.h, int(N), int(octaves), float(base_freq), int(seed),
                                              ctypes.byref(vid)))
        return vid.value

    def submit(self, volume_id, view, out, tag=0):
        r = Request(int(volume_id), view, out.data_ptr(), int(tag))
        check(lib().mrp_server_submit(self.h, ctypes.byref(r)))

    def step(self):
        check(lib().mrp_server_step(self.h))

    def drain(self):
        check(lib().mrp_server_drain(self.h))

    def poll(self, max_n=1024):
        buf = (Completion * max_n)()
        n = ctypes.c_int()
        check(lib().mrp_server_poll(self.h, buf, max_n, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def stats(self):
        st = ServerStats()
        check(lib().mrp_server_get_stats(self.h, ctypes.byref(st)))
        return st

    def is_resident(self, volume_id):
        return bool(lib().mrp_server_is_resident(self.h, int(volume_id)))

    def run_poisson(self, rate_rps, duration_s, sizes, seed=1, buffers_per_size=16):
        spec = LoadSpec()
        spec.rate_rps, spec.duration_s, spec.seed = float(rate_rps), float(duration_s), int(seed)
        spec.n_sizes = len(sizes)
        for i, (w, h) in enumerate(sizes):
            spec.sizes[2 * i], spec.sizes[2 * i + 1] = int(w), int(h)
        spec.buffers_per_size = int(buffers_per_size)
        rep = LoadReport()
        check(lib().mrp_server_run_poisson(self.h, ctypes.byref(spec), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in LoadReport._fields_}
