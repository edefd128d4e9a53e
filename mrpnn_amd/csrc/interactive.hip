// Interactive progressive rendering: VolumeRender::Render(float3* target, Histogram*,
// unsigned* target2, ...) (core/volume.cu:550-628) with its kernels RenderCamera<bool>
// (core/volume.cu:57-196: per-frame radiance, progressive blend, per-pixel histogram),
// Denoise<bool> (core/volume.cu:285-337: checkerboard fill, histogram-similarity 11x11
// filter), ReprojectionDenoise (core/volume.cu:339-358), ClearHis (core/volume.cu:360-366)
// and the tone maps Gamma / ACES (core/render.cuh:1144-1236).
//
// The per-frame radiance comes from the predictor kernels of this library (MRPNN) or the
// path sample shared with the judge (PT); the blend, the histogram update, the filter and the
// tone map are elementwise or stencil work over 144 bytes per pixel.  The filter lists the
// noisy pixels and spreads them over a persistent grid, one pixel's 121 histogram comparisons
// per workgroup, so its cost follows the noisy pixels, not the image.
#include "../../include/mrp_c_api.h"
#include "capi_internal.h"
#include "dev_pathtrace.hpp"
#include "tonemap.hpp"

namespace mrp {

constexpr int kHisto = MRP_HISTO_SIZE;
constexpr int kBins = 3 * kHisto;
static_assert(sizeof(mrp_histogram) == 132, "Histogram is 30 bins + 3 floats");

// Checkerboard pixel set of a frame: RenderCamera renders column j = 2 jj + (i + flip) % 2 of
// row i (core/volume.cu:60-61), Denoise keeps the same pixels at frame 0 (core/volume.cu:295).
__device__ inline bool in_frame_set(int i, int j, int checker, int flip) {
    return !checker || ((((i + flip) % 2) + j) % 2) == 0;
}

// exposure, tone map, saturate, 8 bits per channel, 0xAARRGGBB (core/volume.cu:330-336)
__device__ inline uint32_t tonemap_pack(v3 res, float exposure, int tone) {
    v3 val = res * exposure;
    if (tone == MRP_TONE_GAMMA) val = gamma22(val);
    else if (tone == MRP_TONE_ACES) val = aces(val);
    const uint32_t r = (uint32_t)(255.0f * saturatef_ref(val.x));
    const uint32_t g = (uint32_t)(255.0f * saturatef_ref(val.y));
    const uint32_t b = (uint32_t)(255.0f * saturatef_ref(val.z));
    return 0xff000000u | (r << 16) | (g << 8) | b;
}

__device__ inline v3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
__device__ inline void st3(float* p, v3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// RenderCamera<false> (core/volume.cu:57-131): one path per pixel of the frame's set, blended
// into the target with weight 1 / (1 + n) and counted into the histogram.  Target row i is
// the render row height-1-i (v = (i + jitter) / height).
__global__ __launch_bounds__(256) void k_pt_frame(DevVolume vol, ViewDev vw, int checker, int flip, int fnum,
                                                  int max_bounce, float* __restrict__ target,
                                                  mrp_histogram* __restrict__ histo, float4* __restrict__ rad) {
    const int W = vw.width, H = vw.height;
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= W * H) return;
    const int i = id / W, j = id % W;
    if (!in_frame_set(i, j, checker, flip)) {
        rad[id] = make_float4(0.0f, 0.0f, 0.0f, -2.0f);
        return;
    }
    const int row = H - 1 - i;
    const v3 light = normalize3(vw.light_dir);
    const v3 fwd = normalize3(-vw.ori);
    const float sy = vw.tan_half_fov, sx = vw.tan_half_fov * ((float)W / (float)H);
    unsigned taps = 0;
    const NeeSrc nee = {NEE_TR_MIP0, nullptr, 0};
    v3 res = pt_sample(vol, vw, light, fwd, sx, sy, row, j, (uint32_t)(row * W + j), 0, max_bounce, nee, 1.0f, taps);
    res = mk3(fmaxf(0.0f, res.x), fmaxf(0.0f, res.y), fmaxf(0.0f, res.z));
    rad[id] = make_float4(res.x, res.y, res.z, 0.0f);
    mrp_histogram* h = histo + id;
    if (fnum == 0) {
        for (int k = 0; k < kBins; ++k) h->bin[k] = 0.0f;
        h->total_sample_num = 0.0f;
        h->x2 = 0.0f;
        h->x = 0.0f;
    }
    const float rate = 1.0f / (float)(1 + fnum);
    float* t = target + 3 * (size_t)id;
    st3(t, lerp3(ld3(t), res, rate));
    const v3 r = mk3(res.x / (res.x + 1.0f), res.y / (res.y + 1.0f), res.z / (res.z + 1.0f));
    h->total_sample_num += 1.0f;
    const int bx = (int)(fminf(r.x, 0.999f) * (float)kHisto);
    const int by = (int)(fminf(r.y, 0.999f) * (float)kHisto);
    const int bz = (int)(fminf(r.z, 0.999f) * (float)kHisto);
    h->bin[bx] += 1.0f;
    h->bin[by + kHisto] += 1.0f;
    h->bin[bz + 2 * kHisto] += 1.0f;
    const float l = dot3(r, mk3(1.0f, 1.0f, 1.0f));
    h->x = lerpf_ref(h->x, l, rate);
    h->x2 = lerpf_ref(h->x2, l * l, rate);
}

// RenderCamera<true> (core/volume.cu:133-150): the predictor's max(0, rgb) into the target on
// the frame's set; the predictor image is in render order (row 0 at the top).
__global__ __launch_bounds__(256) void k_commit_predict(const float4* __restrict__ pred, int W, int H, int checker,
                                                        int flip, float* __restrict__ target,
                                                        float4* __restrict__ rad) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= W * H) return;
    const int i = id / W, j = id % W;
    if (!in_frame_set(i, j, checker, flip)) {
        rad[id] = make_float4(0.0f, 0.0f, 0.0f, -2.0f);
        return;
    }
    const float4 p = pred[(size_t)(H - 1 - i) * W + j];
    const v3 res = mk3(fmaxf(0.0f, p.x), fmaxf(0.0f, p.y), fmaxf(0.0f, p.z));
    rad[id] = make_float4(res.x, res.y, res.z, p.w);
    st3(target + 3 * (size_t)id, res);
}

__global__ __launch_bounds__(256) void k_clear_histo(mrp_histogram* __restrict__ histo, int n) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n) return;
    float* h = reinterpret_cast<float*>(histo + id);
    for (int k = 0; k < kBins + 3; ++k) h[k] = 0.0f;
}

// ReprojectionDenoise (core/volume.cu:339-358): the histogram keeps the target's (r, g, b) as
// (total, x, x2), the target is tone-mapped.
__global__ __launch_bounds__(256) void k_reprojection_resolve(const float* __restrict__ target,
                                                              mrp_histogram* __restrict__ histo,
                                                              uint32_t* __restrict__ rgba, int n, float exposure,
                                                              int tone) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n) return;
    const v3 t = ld3(target + 3 * (size_t)id);
    histo[id].total_sample_num = t.x;
    histo[id].x = t.y;
    histo[id].x2 = t.z;
    rgba[id] = tonemap_pack(t, exposure, tone);
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Denoise at frame 0 (checkerboard fill), or without the filter (core/volume.cu:290-327).
__global__ __launch_bounds__(256) void k_resolve_plain(const float* __restrict__ target,
                                                       const mrp_histogram* __restrict__ histo,
                                                       uint32_t* __restrict__ rgba, int W, int H, int frame0, int flip,
                                                       float exposure, int tone) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= W * H) return;
    const int r = id / W, c = id % W;
    v3 res;
    if (frame0 && ((((r + flip) % 2) + c) % 2) != 0) {
        // UnRoll clamps the row to [0, H-1] and the column to [0, W-1] (core/volume.cu:230-234)
        const v3 a = ld3(target + 3 * (size_t)(clampi(r - 1, 0, H - 1) * W + c));
        const v3 b = ld3(target + 3 * (size_t)(clampi(r + 1, 0, H - 1) * W + c));
        const v3 d = ld3(target + 3 * (size_t)(r * W + clampi(c - 1, 0, W - 1)));
        const v3 e = ld3(target + 3 * (size_t)(r * W + clampi(c + 1, 0, W - 1)));
        res = (a + b + d + e) / 4.0f;
    } else {
        res = ld3(target + 3 * (size_t)id);
    }
    (void)histo;
    rgba[id] = tonemap_pack(res, exposure, tone);
}

// Denoise<true> after frame 0 (core/volume.cu:304-327): where the luminance variance of a pixel
// exceeds 0.01, the average of its 11x11 neighbourhood (clamped at the borders) weighted by
// max(0, 1 - 1.2 Compare(centre, neighbour)) (Compare: core/volume.cu:236-255).
//
// Noisy pixels are few and clustered, so a pass over all pixels tone-maps the quiet ones and
// lists the noisy ones; a persistent grid then takes the listed pixels one per 128-thread
// workgroup: 121 threads each compare one neighbour's histogram with the centre's (30 bins,
// read through L1 from the 132-byte records) and stage its weight and target in LDS, and one
// thread sums them in the reference's order, so the result is the same float for float.
constexpr int kDR = 5, kDW = 2 * kDR + 1, kDN = kDW * kDW;

__global__ __launch_bounds__(256) void k_denoise_mark(const float* __restrict__ target,
                                                      const mrp_histogram* __restrict__ histo,
                                                      uint32_t* __restrict__ rgba, int n, float exposure, int tone,
                                                      int* __restrict__ list, unsigned* __restrict__ count) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    bool noisy = false;
    if (id < n) {
        const float hx2 = histo[id].x2, hxm = histo[id].x;
        noisy = fabsf(hx2 - hxm * hxm) > 0.01f;
        if (!noisy) rgba[id] = tonemap_pack(ld3(target + 3 * (size_t)id), exposure, tone);
    }
    const unsigned long long mask = __ballot(noisy);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    unsigned base = 0;
    if (lane == __ffsll((long long)mask) - 1) base = atomicAdd(count, (unsigned)__popcll(mask));
    base = __shfl(base, __ffsll((long long)mask) - 1);
    if (noisy) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = id;
}

__global__ __launch_bounds__(128) void k_denoise_pixels(const float* __restrict__ target,
                                                        const mrp_histogram* __restrict__ histo,
                                                        uint32_t* __restrict__ rgba, int W, int H, float exposure,
                                                        int tone, const int* __restrict__ list,
                                                        const unsigned* __restrict__ count) {
    __shared__ float s_w[kDN];
    __shared__ float s_t[3][kDN];
    const int tid = threadIdx.x;
    const int n = (int)*count;
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int id = list[q];
        const int r = id / W, c = id % W;
        if (tid < kDN) {
            const int nr = clampi(r + tid / kDW - kDR, 0, H - 1), nc = clampi(c + tid % kDW - kDR, 0, W - 1);
            const size_t nid = (size_t)nr * W + nc;
            const float* hx = reinterpret_cast<const float*>(histo + id);
            const float* hy = reinterpret_cast<const float*>(histo + nid);
            float bx[kBins], by[kBins];
            for (int b = 0; b < kBins; ++b) {
                bx[b] = hx[b];
                by[b] = hy[b];
            }
            const float nx = hx[kBins], ny = hy[kBins];
            const float sqrt_y_x = sqrtf(ny / nx), sqrt_x_y = sqrtf(nx / ny);
            float p = 0.0f, sum = 0.0f;
            for (int b = 0; b < kBins; ++b) {  // the skip of empty bin pairs as selects of exact zeros
                const bool used = bx[b] != 0.0f || by[b] != 0.0f;
                const float t = sqrt_y_x * bx[b] - sqrt_x_y * by[b];
                const float qv = t * t / (bx[b] + by[b]);
                p += used ? 1.0f : 0.0f;
                sum += used ? qv : 0.0f;
            }
            const float cmp = p == 0.0f ? 0.0f : sum / p;
            s_w[tid] = fmaxf(0.0f, 1.0f - 1.2f * cmp);
            s_t[0][tid] = target[3 * nid];
            s_t[1][tid] = target[3 * nid + 1];
            s_t[2][tid] = target[3 * nid + 2];
        }
        __syncthreads();
        if (tid == 0) {
            v3 acc = mk3(0.0f, 0.0f, 0.0f);
            float ws = 0.0f;
            for (int nb = 0; nb < kDN; ++nb) {  // i outer, j inner as the reference
                const float w = s_w[nb];
                acc = acc + mk3(s_t[0][nb], s_t[1][nb], s_t[2][nb]) * w;
                ws += w;
            }
            rgba[id] = tonemap_pack(acc / ws, exposure, tone);
        }
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

}  // namespace mrp

using namespace mrp;

struct mrp_interactive {
    mrp_ctx* ctx = nullptr;
    int W = 0, H = 0;
    int checkerboard = 1;    // VolumeRender::checkboard (core/volume.hpp:50)
    bool last_predict = true;  // core/volume.hpp:51
    int flip = 0;            // flip_cpu (core/volume.cu:481)
    float exposure = 1.0f;   // the device global exposure (core/volume.cu:52)
    float4* d_rad = nullptr;   // the frame's radiance, target layout
    float4* d_pred = nullptr;  // predictor output, render layout
    int* d_list = nullptr;        // noisy pixels of the filter
    unsigned* d_count = nullptr;  // their number
    int filter_grid = 0;          // persistent workgroups of the filter: 16 per CU
};

extern "C" int mrp_interactive_create(mrp_ctx* c, int width, int height, mrp_interactive** out) {
    if (!c || !out) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (width < 1 || height < 1 || (int64_t)width * height > (1 << 28))
        return mrp_fail(MRP_ERR_INVALID, "interactive size %d x %d out of range", width, height);
    mrp_interactive* it = new mrp_interactive();
    it->ctx = c;
    it->W = width;
    it->H = height;
    const size_t bytes = (size_t)width * height * sizeof(float4);
    if (hipMalloc(&it->d_rad, bytes) != hipSuccess || hipMalloc(&it->d_pred, bytes) != hipSuccess ||
        hipMalloc(&it->d_list, (size_t)width * height * sizeof(int)) != hipSuccess ||
        hipMalloc(&it->d_count, sizeof(unsigned)) != hipSuccess) {
        (void)hipFree(it->d_rad);
        (void)hipFree(it->d_pred);
        (void)hipFree(it->d_list);
        (void)hipFree(it->d_count);
        delete it;
        return mrp_fail(MRP_ERR_NOMEM, "interactive buffers (%zu bytes)", 2 * bytes);
    }
    it->filter_grid = 16 * (c->props.multiProcessorCount > 0 ? c->props.multiProcessorCount : 256);
    *out = it;
    return MRP_OK;
}

extern "C" void mrp_interactive_destroy(mrp_interactive* it) {
    if (!it) return;
    (void)hipFree(it->d_rad);
    (void)hipFree(it->d_pred);
    (void)hipFree(it->d_list);
    (void)hipFree(it->d_count);
    delete it;
}

extern "C" int mrp_interactive_set_checkerboard(mrp_interactive* it, int on) {
    if (!it) return mrp_fail(MRP_ERR_INVALID, "null argument");
    it->checkerboard = on ? 1 : 0;
    return MRP_OK;
}

extern "C" int mrp_interactive_set_exposure(mrp_interactive* it, float exposure) {
    if (!it) return mrp_fail(MRP_ERR_INVALID, "null argument");
    it->exposure = exposure;
    return MRP_OK;
}

extern "C" int mrp_interactive_radiance(mrp_interactive* it, float* dev_out4) {
    if (!it || !dev_out4) return mrp_fail(MRP_ERR_INVALID, "null argument");
    hipError_t e = hipMemcpyAsync(dev_out4, it->d_rad, (size_t)it->W * it->H * sizeof(float4),
                                  hipMemcpyDeviceToDevice, it->ctx->stream);
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "radiance copy: %s", hipGetErrorString(e));
    return MRP_OK;
}

extern "C" int mrp_render_interactive(mrp_interactive* it, mrp_volume* v, const mrp_view* vw, int render_type,
                                      int precision, int max_bounce, int randseed, int tone_type, int denoise,
                                      float* dev_target, mrp_histogram* dev_histo, uint32_t* dev_rgba) {
    if (!it || !v || !vw || !dev_target || !dev_histo || !dev_rgba) return mrp_fail(MRP_ERR_INVALID, "null argument");
    mrp_ctx* c = it->ctx;
    if (v->ctx != c) return mrp_fail(MRP_ERR_INVALID, "volume belongs to another context");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (vw->width != it->W || vw->height != it->H)
        return mrp_fail(MRP_ERR_INVALID, "view is %d x %d, the renderer %d x %d", vw->width, vw->height, it->W, it->H);
    if (render_type != MRP_RT_PT && render_type != MRP_RT_MRPNN && render_type != MRP_RT_RPNN)
        return mrp_fail(MRP_ERR_INVALID, "unknown render type %d", render_type);
    if (tone_type < 0 || tone_type > 2) return mrp_fail(MRP_ERR_INVALID, "unknown tone type %d", tone_type);
    if (randseed < 0) return mrp_fail(MRP_ERR_INVALID, "randseed must be >= 0");
    const bool predict = render_type != MRP_RT_PT;
    if (!predict && max_bounce < 1) return mrp_fail(MRP_ERR_INVALID, "max_bounce >= 1 required");
    if (predict && (precision < MRP_FP32 || precision > MRP_BF16_1T))
        return mrp_fail(MRP_ERR_INVALID, "unknown precision %d", precision);
    if (render_type == MRP_RT_RPNN && precision != MRP_FP32 && precision != MRP_BF16)  // as mrp_render_rpnn
        return mrp_fail(MRP_ERR_INVALID, "RPNN runs in MRP_FP32 or MRP_BF16, not precision %d", precision);
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before rendering");

    mrp_view view = *vw;
    view.frame = (uint32_t)randseed;  // the jitter and path streams of this frame
    view.volume_index = 0;
    const int W = it->W, H = it->H, n = W * H;
    const int flip = it->flip;
    it->flip = (it->flip + 1) % 2;  // core/volume.cu:578
    const int blocks = (n + 255) / 256;
    hipError_t e = hipSuccess;
    if (predict && !it->last_predict) {  // ClearHis when switching from PT (core/volume.cu:601-604)
        hipLaunchKernelGGL(k_clear_histo, dim3(blocks), dim3(256), 0, c->stream, dev_histo, n);
        e = hipGetLastError();
    }
    it->last_predict = predict;
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "interactive: %s", hipGetErrorString(e));
    if (predict) {
        // the batched predictor (its FP32 form equals the plain per-frame kernel bit for bit),
        // marching only this frame's checkerboard half: target row i is render row H-1-i, so
        // the set ((i + flip) + j) even is (row + col + H - 1 + flip) even
        float* pred = reinterpret_cast<float*>(it->d_pred);
        mrp_volume* vols[1] = {v};
        const int checker = it->checkerboard ? 1 + ((H - 1 + flip) & 1) : 0;
        if (render_type == MRP_RT_RPNN) {  // RenderCamera<true, RPNN>: FP32, or the BF16 network
            rc = mrp_render_rpnn_checker(c, v, &view, precision, pred, checker);
        } else {
            rc = mrp_prepare_views(c, vols, 1, &view, 1);
            if (!rc) rc = mrp_render_batch_checker(c, vols, 1, &view, 1, precision, &pred, &checker);
        }
        if (rc) return rc;
        hipLaunchKernelGGL(k_commit_predict, dim3(blocks), dim3(256), 0, c->stream, it->d_pred, W, H,
                           it->checkerboard, flip, dev_target, it->d_rad);
        e = hipGetLastError();
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_reprojection_resolve, dim3(blocks), dim3(256), 0, c->stream, dev_target, dev_histo,
                               dev_rgba, n, it->exposure, tone_type);
            e = hipGetLastError();
        }
    } else {
        // baked TR for next-event estimation, as CalculateRadiance (core/render.cuh:589)
        rc = mrp_prepare_view(c, v, &view);
        if (rc) return rc;
        const int fnum = it->checkerboard ? randseed / 2 : randseed;
        hipLaunchKernelGGL(k_pt_frame, dim3(blocks), dim3(256), 0, c->stream, mrp_dev_volume(v), mrp_to_dev(&view),
                           it->checkerboard, flip, fnum, max_bounce, dev_target, dev_histo, it->d_rad);
        e = hipGetLastError();
        if (e == hipSuccess) {
            if (denoise && randseed != 0) {
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, c->stream, it->d_count);
                hipLaunchKernelGGL(k_denoise_mark, dim3(blocks), dim3(256), 0, c->stream, dev_target, dev_histo,
                                   dev_rgba, n, it->exposure, tone_type, it->d_list, it->d_count);
                hipLaunchKernelGGL(k_denoise_pixels, dim3(it->filter_grid), dim3(128), 0, c->stream, dev_target,
                                   dev_histo, dev_rgba, W, H, it->exposure, tone_type, it->d_list, it->d_count);
            } else {
                hipLaunchKernelGGL(k_resolve_plain, dim3(blocks), dim3(256), 0, c->stream, dev_target, dev_histo,
                                   dev_rgba, W, H, randseed == 0 ? 1 : 0, flip, it->exposure, tone_type);
            }
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "interactive: %s", hipGetErrorString(e));
    return MRP_OK;
}
