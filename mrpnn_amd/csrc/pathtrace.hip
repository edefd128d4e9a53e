// Delta-tracking path tracer: the correctness judge of the predictor.
//
// RenderCamera<false> -> CalculateRadiance (core/volume.cu:198-228, core/render.cuh:535-597)
// with Woodcock tracking DeterminateNextVertex (core/render.cuh:399-434) and next-event
// estimation either through the stochastic Tr (ratio tracking to 0.5, then delta tracking,
// core/render.cuh:60-103; unbiased), the TR mip 0 of Fill_TR, or the cached stochastic bake
// of Fill_BakeTR that the reference's path tracer reads (core/render.cuh:589 after
// Update_CachedTR, core/volume.cu:1669-1696).  Every (pixel, sample) path draws from its own
// Philox stream {seed, frame} / {pixel, RNG_PT, sample, *}; per-pixel mean and M2 accumulate
// in FP64 (Welford), so a render can be split into sample chunks and merged exactly.  Also
// here: the point-radiance paths of GetSamples (the training targets of dataGen.cu) and the
// transmittance probes of GetTr / GetTrs.
#include <math.h>
#include <string.h>

#include "capi_internal.h"
#include "dev_pathtrace.hpp"

namespace mrp {

__global__ __launch_bounds__(64) void k_pathtrace(DevVolume vol, ViewDev vw, int spp0, int spp, int max_bounce,
                                                  NeeSrc nee, float tr_scale, double* __restrict__ mean3,
                                                  double* __restrict__ m2_3, unsigned long long* __restrict__ taps_total) {
    const int col = blockIdx.x * 8 + (threadIdx.x & 7);
    const int row = blockIdx.y * 8 + (threadIdx.x >> 3);
    unsigned taps = 0;
    if (row < vw.height && col < vw.width) {
        const uint32_t pix = (uint32_t)(row * vw.width + col);
        const size_t p = (size_t)pix;
        double mean[3], m2[3];
        for (int c = 0; c < 3; ++c) {
            mean[c] = spp0 ? mean3[3 * p + c] : 0.0;
            m2[c] = spp0 ? m2_3[3 * p + c] : 0.0;
        }
        const v3 light = normalize3(vw.light_dir);
        const v3 fwd = normalize3(-vw.ori);
        const float sy = vw.tan_half_fov, sx = vw.tan_half_fov * ((float)vw.width / (float)vw.height);
        for (int s = spp0; s < spp0 + spp; ++s) {
            const v3 res = pt_sample(vol, vw, light, fwd, sx, sy, row, col, pix, s, max_bounce, nee, tr_scale, taps);
            const double x[3] = {(double)res.x, (double)res.y, (double)res.z};
            for (int c = 0; c < 3; ++c) {
                double d = x[c] - mean[c];
                mean[c] += d / (double)(s + 1);
                m2[c] += d * (x[c] - mean[c]);
            }
        }
        for (int c = 0; c < 3; ++c) {
            mean3[3 * p + c] = mean[c];
            m2_3[3 * p + c] = m2[c];
        }
    }
    unsigned long long t = taps;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(taps_total, t);
}

// Fill_BakeTR (core/volume.cu:1562-1587): texel (x, y, z) of a res^3 grid over the unit box holds
// the mean of 5 stochastic Tr estimates towards the light from its centre where the density
// there is positive, else of 2 (TR_MUL = 1).  Each estimate draws from its own Philox stream
// {seed, 1} / {texel, RNG_BAKE, estimate}.
__global__ __launch_bounds__(256) void k_bake_cached_tr(DevVolume vol, int res, v3 light, float alpha, float tr_scale,
                                                        uint32_t seed, float* __restrict__ out) {
    const size_t total = (size_t)res * res * res;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const int z = (int)(i % res), y = (int)((i / res) % res), x = (int)(i / ((size_t)res * res));
        const v3 pos = (mk3((float)x, (float)y, (float)z) + 0.5f) / (float)res - 0.5f;
        const int k = density_full(vol, pos) > 0.0f ? 5 : 2;
        const float dis = ray_box_distance(pos, light);
        unsigned taps = 0;
        float acc = 0.0f;
        for (int e = 0; e < k; ++e) {
            PhiloxStream rs(seed, 1u, (uint32_t)i, RNG_BAKE, (uint32_t)e);
            acc += tr_stochastic(vol, rs, pos, light, dis, alpha, tr_scale, taps);
        }
        out[i] = acc / (float)k;
    }
}

// GetSampleMulti (core/volume.cu:36-47): task i's point-radiance estimate over spp paths, each
// from its own Philox stream {seed, 0} / {task, RNG_POINT, sample}; FP64 mean and M2 (Welford).
__global__ __launch_bounds__(64) void k_point_samples(DevVolume vol, int n, const float* __restrict__ ori3,
                                                      const float* __restrict__ dir3, const float* __restrict__ light3,
                                                      const float* __restrict__ alpha, const float* __restrict__ g,
                                                      const float* __restrict__ scatter, v3 light_color, int max_bounce,
                                                      int spp, uint32_t seed, float tr_scale, float* __restrict__ out3,
                                                      double* __restrict__ m2_3) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const v3 ori = mk3(ori3[3 * i], ori3[3 * i + 1], ori3[3 * i + 2]);
    const v3 dir = mk3(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
    const v3 light = mk3(light3[3 * i], light3[3 * i + 1], light3[3 * i + 2]);
    double mean[3] = {0.0, 0.0, 0.0}, m2[3] = {0.0, 0.0, 0.0};
    unsigned taps = 0;
    for (int s = 0; s < spp; ++s) {
        PhiloxStream rs(seed, 0u, (uint32_t)i, RNG_POINT, (uint32_t)s);
        const v3 r = point_sample(vol, rs, ori, dir, light, light_color, scatter[i], alpha[i], g[i], max_bounce, tr_scale,
                                  taps);
        const double x[3] = {(double)r.x, (double)r.y, (double)r.z};
        for (int c = 0; c < 3; ++c) {
            const double d = x[c] - mean[c];
            mean[c] += d / (double)(s + 1);
            m2[c] += d * (x[c] - mean[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        out3[3 * (size_t)i + c] = (float)mean[c];
        if (m2_3) m2_3[3 * (size_t)i + c] = m2[c];
    }
}

// ShadowTerm_TR(s) (core/render.cuh:656-680): task i's mean over spp of Tr(ori -> light) times
// HG(dot(dir, light), g), each estimate from its own Philox stream {seed, 2} / {task, RNG_TR, sample}.
__global__ __launch_bounds__(64) void k_tr_samples(DevVolume vol, int n, const float* __restrict__ ori3,
                                                   const float* __restrict__ dir3, v3 light, float alpha, float g,
                                                   int spp, uint32_t seed, float tr_scale, float* __restrict__ out) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const v3 ori = mk3(ori3[3 * i], ori3[3 * i + 1], ori3[3 * i + 2]);
    const v3 dir = mk3(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
    const float dis = ray_box_distance(ori, light);
    const float phase = henyey_greenstein(dot3(dir, light), g);
    unsigned taps = 0;
    double acc = 0.0;
    for (int s = 0; s < spp; ++s) {
        PhiloxStream rs(seed, 2u, (uint32_t)i, RNG_TR, (uint32_t)s);
        acc += (double)(tr_stochastic(vol, rs, ori, light, dis, alpha, tr_scale, taps) * phase);
    }
    out[i] = (float)(acc / (double)spp);
}

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_volume_bake_cached_tr(mrp_volume* v, const float light_dir[3], float alpha, float tr_scale,
                                         uint32_t seed, int res) {
    if (!v || !light_dir) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (!(alpha > 0.0f) || !isfinite(alpha) || !(tr_scale > 0.0f) || !isfinite(tr_scale))
        return mrp_fail(MRP_ERR_INVALID, "alpha and tr_scale must be positive and finite");
    if (res == 0) res = 768;
    if (res < 2 || res > 1024) return mrp_fail(MRP_ERR_INVALID, "res in [2, 1024] (0: 768)");
    const float l2 = light_dir[0] * light_dir[0] + light_dir[1] * light_dir[1] + light_dir[2] * light_dir[2];
    if (!(l2 > 0.0f) || !isfinite(l2)) return mrp_fail(MRP_ERR_INVALID, "light_dir must be a finite nonzero vector");
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before the bake");
    if (v->ctr_valid && v->ctr_res == res && v->ctr_alpha == alpha && v->ctr_scale == tr_scale && v->ctr_seed == seed &&
        v->ctr_light[0] == light_dir[0] && v->ctr_light[1] == light_dir[1] && v->ctr_light[2] == light_dir[2])
        return MRP_OK;  // cached on the exact key, as Update_CachedTR (core/volume.cu:1672-1674)
    mrp_ctx* c = v->ctx;
    if (v->ctr_res != res) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "stream sync failed");
        hipFree(v->d_ctr);
        v->d_ctr = nullptr;
        v->ctr_res = 0;
        v->ctr_valid = false;
        const size_t bytes = (size_t)res * res * res * sizeof(float);
        if (hipMalloc(&v->d_ctr, bytes) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "cached TR allocation of %zu bytes failed", bytes);
        v->ctr_res = res;
    }
    v->ctr_valid = false;
    const size_t total = (size_t)res * res * res;
    const int grid = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_bake_cached_tr, dim3(grid), dim3(256), 0, c->stream, mrp_dev_volume(v), res,
                       normalize3(mk3(light_dir[0], light_dir[1], light_dir[2])), alpha, tr_scale, seed, v->d_ctr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "cached TR bake: %s", hipGetErrorString(e));
    memcpy(v->ctr_light, light_dir, sizeof v->ctr_light);
    v->ctr_alpha = alpha;
    v->ctr_scale = tr_scale;
    v->ctr_seed = seed;
    v->ctr_valid = true;
    return MRP_OK;
}

extern "C" int mrp_volume_download_cached_tr(mrp_volume* v, float* host) {
    if (!v || !host) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (!v->ctr_valid) return mrp_fail(MRP_ERR_STATE, "no cached TR bake");
    const size_t bytes = (size_t)v->ctr_res * v->ctr_res * v->ctr_res * sizeof(float);
    if (hipMemcpyAsync(host, v->d_ctr, bytes, hipMemcpyDeviceToHost, v->ctx->stream) != hipSuccess ||
        hipStreamSynchronize(v->ctx->stream) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "download failed");
    return MRP_OK;
}

extern "C" int mrp_point_radiance(mrp_ctx* c, mrp_volume* v, int n, const float* dev_ori3, const float* dev_dir3,
                                  const float* dev_light3, const float* dev_alpha, const float* dev_g,
                                  const float* dev_scatter, const float light_color[3], int multi_scatter, int spp,
                                  uint32_t seed, float tr_scale, float* dev_out3, double* dev_m2_3) {
    if (!c || !v || !light_color) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (n < 0 || spp < 1 || multi_scatter < 1 || !(tr_scale > 0.0f))
        return mrp_fail(MRP_ERR_INVALID, "n >= 0, spp >= 1, multi_scatter >= 1, tr_scale > 0 required");
    if (n == 0) return MRP_OK;
    if (!dev_ori3 || !dev_dir3 || !dev_light3 || !dev_alpha || !dev_g || !dev_scatter || !dev_out3)
        return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run first");
    hipLaunchKernelGGL(k_point_samples, dim3((n + 63) / 64), dim3(64), 0, c->stream, mrp_dev_volume(v), n, dev_ori3,
                       dev_dir3, dev_light3, dev_alpha, dev_g, dev_scatter,
                       mk3(light_color[0], light_color[1], light_color[2]), multi_scatter, spp, seed, tr_scale, dev_out3,
                       dev_m2_3);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "point radiance: %s", hipGetErrorString(e));
    return MRP_OK;
}

extern "C" int mrp_tr_samples(mrp_ctx* c, mrp_volume* v, int n, const float* dev_ori3, const float* dev_dir3,
                              const float light_dir[3], float alpha, float g, int spp, uint32_t seed, float tr_scale,
                              float* dev_out) {
    if (!c || !v || !light_dir) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (n < 0 || spp < 1 || !(alpha > 0.0f) || !(tr_scale > 0.0f))
        return mrp_fail(MRP_ERR_INVALID, "n >= 0, spp >= 1, alpha > 0, tr_scale > 0 required");
    if (n == 0) return MRP_OK;
    if (!dev_ori3 || !dev_dir3 || !dev_out) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run first");
    hipLaunchKernelGGL(k_tr_samples, dim3((n + 63) / 64), dim3(64), 0, c->stream, mrp_dev_volume(v), n, dev_ori3,
                       dev_dir3, normalize3(mk3(light_dir[0], light_dir[1], light_dir[2])), alpha, g, spp, seed,
                       tr_scale, dev_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "transmittance samples: %s", hipGetErrorString(e));
    return MRP_OK;
}

extern "C" int mrp_pt_render(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int spp0, int spp, int max_bounce,
                             int nee_baked, float tr_scale, double* dev_mean3, double* dev_m2_3) {
    if (!c || !v || !dev_mean3 || !dev_m2_3) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (spp0 < 0 || spp < 1 || max_bounce < 1 || !(tr_scale > 0.0f))
        return mrp_fail(MRP_ERR_INVALID, "spp0 >= 0, spp >= 1, max_bounce >= 1, tr_scale > 0 required");
    if (nee_baked < 0 || nee_baked > 2) return mrp_fail(MRP_ERR_INVALID, "nee_baked is 0, 1 or 2");
    if (nee_baked == 1 && !(v->tr_valid && v->tr_alpha == vw->alpha && v->tr_light[0] == vw->light_dir[0] &&
                            v->tr_light[1] == vw->light_dir[1] && v->tr_light[2] == vw->light_dir[2]))
        return mrp_fail(MRP_ERR_STATE, "baked NEE needs TR mips for this light and alpha");
    if (nee_baked == 2 && !(v->ctr_valid && v->ctr_alpha == vw->alpha && v->ctr_scale == tr_scale &&
                            v->ctr_light[0] == vw->light_dir[0] && v->ctr_light[1] == vw->light_dir[1] &&
                            v->ctr_light[2] == vw->light_dir[2]))
        return mrp_fail(MRP_ERR_STATE, "cached NEE needs mrp_volume_bake_cached_tr for this light, alpha and tr_scale");
    const NeeSrc nee = {nee_baked, nee_baked == 2 ? v->d_ctr : nullptr, nee_baked == 2 ? v->ctr_res : 0};
    unsigned long long* d_taps = nullptr;
    if (hipMalloc(&d_taps, sizeof(unsigned long long)) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "allocation failed");
    hipError_t e = hipMemsetAsync(d_taps, 0, sizeof(unsigned long long), c->stream);
    dim3 grid((vw->width + 7) / 8, (vw->height + 7) / 8);
    if (e == hipSuccess) {
        if (c->timing) hipEventRecord(c->ev[0], c->stream);
        hipLaunchKernelGGL(k_pathtrace, grid, dim3(64), 0, c->stream, mrp_dev_volume(v), mrp_to_dev(vw), spp0, spp,
                           max_bounce, nee, tr_scale, dev_mean3, dev_m2_3, d_taps);
        e = hipGetLastError();
        if (c->timing) hipEventRecord(c->ev[1], c->stream);
    }
    unsigned long long taps = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&taps, d_taps, sizeof taps, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d_taps);
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "path tracer: %s", hipGetErrorString(e));
    c->stats.pixels = (int64_t)vw->width * vw->height;
    c->stats.march_taps = (int64_t)taps;
    c->stats.active_pixels = -1;
    if (c->timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        c->stats.ms_total = ms;
        c->stats.ms_predict = ms;
        c->stats.ms_march = 0;
    }
    return MRP_OK;
}
