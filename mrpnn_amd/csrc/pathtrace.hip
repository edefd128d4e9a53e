// Delta-tracking path tracer: the correctness judge of the predictor.
//
// RenderCamera<false> -> CalculateRadiance (core/volume.cu:198-228, core/render.cuh:535-597)
// with Woodcock tracking DeterminateNextVertex (core/render.cuh:399-434) and next-event
// estimation either through the stochastic Tr (ratio tracking to 0.5, then delta tracking,
// core/render.cuh:60-103; unbiased) or the baked TR mip 0 as the reference does
// (core/render.cuh:589).  Every (pixel, sample) path draws from its own Philox stream
// {seed, frame} / {pixel, RNG_PT, sample, *}; per-pixel mean and M2 accumulate in FP64
// (Welford), so a render can be split into sample chunks and merged exactly.
#include "capi_internal.h"
#include "dev_pathtrace.hpp"

namespace mrp {

__global__ __launch_bounds__(64) void k_pathtrace(DevVolume vol, ViewDev vw, int spp0, int spp, int max_bounce,
                                                  int nee_baked, float tr_scale, double* __restrict__ mean3,
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
            const v3 res = pt_sample(vol, vw, light, fwd, sx, sy, row, col, pix, s, max_bounce, nee_baked, tr_scale, taps);
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

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_pt_render(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int spp0, int spp, int max_bounce,
                             int nee_baked, float tr_scale, double* dev_mean3, double* dev_m2_3) {
    if (!c || !v || !dev_mean3 || !dev_m2_3) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (spp0 < 0 || spp < 1 || max_bounce < 1 || !(tr_scale > 0.0f))
        return mrp_fail(MRP_ERR_INVALID, "spp0 >= 0, spp >= 1, max_bounce >= 1, tr_scale > 0 required");
    if (nee_baked && !(v->tr_valid && v->tr_alpha == vw->alpha && v->tr_light[0] == vw->light_dir[0] &&
                       v->tr_light[1] == vw->light_dir[1] && v->tr_light[2] == vw->light_dir[2]))
        return mrp_fail(MRP_ERR_STATE, "baked NEE needs TR mips for this light and alpha");
    unsigned long long* d_taps = nullptr;
    if (hipMalloc(&d_taps, sizeof(unsigned long long)) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "allocation failed");
    hipError_t e = hipMemsetAsync(d_taps, 0, sizeof(unsigned long long), c->stream);
    dim3 grid((vw->width + 7) / 8, (vw->height + 7) / 8);
    if (e == hipSuccess) {
        if (c->timing) hipEventRecord(c->ev[0], c->stream);
        hipLaunchKernelGGL(k_pathtrace, grid, dim3(64), 0, c->stream, mrp_dev_volume(v), mrp_to_dev(vw), spp0, spp,
                           max_bounce, nee_baked, tr_scale, dev_mean3, dev_m2_3, d_taps);
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
