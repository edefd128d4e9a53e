// Plain per-frame FP32 predictor: the drop-in baseline.
//
// One thread per pixel, one launch per frame, in the structure of RenderCamera<true,MRPNN>
// (core/volume.cu:198-228) -> NNPredict<MRPNN> (core/render.cuh:808-875): ray, first-hit
// march, basis, 192-point gather into thread-private arrays, scalar FP32 MLP, composition,
// progressive average.  8x8 blocks = one 64-lane wavefront per block (the reference's 8x4
// blocks are one 32-lane warp).
#include "dev_common.hpp"
#include "dev_mlp_fp32.hpp"
#include "kernels.h"

namespace mrp {

__global__ __launch_bounds__(64) void k_predict_plain(DevVolume vol, const StencilEntry* __restrict__ stencil,
                                                      const float* __restrict__ W, ViewDev vw, float4* __restrict__ out,
                                                      float lerp_rate, int first_frame) {
    const int col = blockIdx.x * 8 + (threadIdx.x & 7);
    const int row = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (row >= vw.height || col >= vw.width) return;
    PixelState st = pixel_state(vol, vw, row, col);
    float4 res = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // sky, distance -1 -> 0 (volume.cu:224)
    if (!st.active && vol.env) {  // SkyBox(dir) (render.cuh:829,874)
        const v3 sky = sky_box(vol, normalize3(pixel_ray(vw, row, col, (uint32_t)(row * vw.width + col))));
        res = make_float4(fmaxf(0.0f, sky.x), fmaxf(0.0f, sky.y), fmaxf(0.0f, sky.z), 0.0f);
    }
    if (st.active) {
        float X[kPC], Hh[kPC], S[kPC];
        for (int i = 0; i < kPC; ++i) {
            v3 f = gather_point(vol, stencil[i], i, st, vw.g, vw.alpha);
            X[i] = f.x;
            Hh[i] = f.y;
            S[i] = f.z;
        }
        Fp32Mlp m;
        m.W = W;
        m.X = X;
        m.S = S;
        m.H = Hh;
        v3 sr = vw.scatter_rate / 1.001f;  // render.cuh:842
        v3 rgb = m.run(gamma_of(st), vw.g, sr);
        v3 r = (rgb * vw.light_color) * vw.scatter_rate;  // render.cuh:874
        res = make_float4(fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z), fmaxf(0.0f, st.dist));
    }
    const size_t idx = (size_t)row * vw.width + col;
    if (first_frame) {
        out[idx] = res;
    } else {  // lerp(result, res, 1/(1+frameNum)) (volume.cu:226-227)
        float4 o = out[idx];
        out[idx] = make_float4(lerp_ref(o.x, res.x, lerp_rate), lerp_ref(o.y, res.y, lerp_rate),
                               lerp_ref(o.z, res.z, lerp_rate), lerp_ref(o.w, res.w, lerp_rate));
    }
}

// Features exactly as the plain kernel computes them, for parity and tooling.
__global__ __launch_bounds__(64) void k_gather_features(DevVolume vol, const StencilEntry* __restrict__ stencil,
                                                        ViewDev vw, float* __restrict__ feats, int* __restrict__ active,
                                                        float* __restrict__ pos3, int* __restrict__ taps) {
    const int col = blockIdx.x * 8 + (threadIdx.x & 7);
    const int row = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (row >= vw.height || col >= vw.width) return;
    PixelState st = pixel_state(vol, vw, row, col);
    const size_t p = (size_t)row * vw.width + col;
    float* f = feats + p * kNFEAT;
    for (int i = 0; i < kPC; ++i) {
        v3 x = st.active ? gather_point(vol, stencil[i], i, st, vw.g, vw.alpha) : mk3(0.0f, 0.0f, 0.0f);
        f[i] = x.x;
        f[kPC + i] = x.y;
        f[2 * kPC + i] = x.z;
    }
    f[3 * kPC] = gamma_of(st);
    if (active) active[p] = st.active;
    if (pos3) {
        pos3[3 * p] = st.pos.x;
        pos3[3 * p + 1] = st.pos.y;
        pos3[3 * p + 2] = st.pos.z;
    }
    if (taps) taps[p] = st.taps;
}

// RadiancePredictInternal on feature rows [X_Val | X_Val_Hg | X_Val_Sub | gamma].
__global__ __launch_bounds__(64) void k_predict_features_fp32(const float* __restrict__ feats, int n,
                                                              const float* __restrict__ W, v3 sr, float g,
                                                              float* __restrict__ rgb) {
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= n) return;
    const float* f = feats + (size_t)r * kNFEAT;
    Fp32Mlp m;
    m.W = W;
    m.X = f;
    m.H = f + kPC;
    m.S = f + 2 * kPC;
    v3 o = m.run(f[3 * kPC], g, sr);
    rgb[3 * (size_t)r] = o.x;
    rgb[3 * (size_t)r + 1] = o.y;
    rgb[3 * (size_t)r + 2] = o.z;
}

// ------------------------------------------------------------------------------------------
hipError_t launch_predict_plain(const DevVolume& vol, const StencilEntry* stencil, const float* W, const ViewDev& vw,
                                float* out4, float lerp_rate, int first_frame, hipStream_t s) {
    dim3 grid((vw.width + 7) / 8, (vw.height + 7) / 8);
    hipLaunchKernelGGL(k_predict_plain, grid, dim3(64), 0, s, vol, stencil, W, vw, (float4*)out4, lerp_rate,
                       first_frame);
    return hipGetLastError();
}
hipError_t launch_gather_features(const DevVolume& vol, const StencilEntry* stencil, const ViewDev& vw, float* feats,
                                  int* active, float* pos3, int* taps, hipStream_t s) {
    dim3 grid((vw.width + 7) / 8, (vw.height + 7) / 8);
    hipLaunchKernelGGL(k_gather_features, grid, dim3(64), 0, s, vol, stencil, vw, feats, active, pos3, taps);
    return hipGetLastError();
}
hipError_t launch_predict_features_fp32(const float* feats, int n, const float* W, v3 sr, float g, float* rgb,
                                        hipStream_t s) {
    hipLaunchKernelGGL(k_predict_features_fp32, dim3((n + 63) / 64), dim3(64), 0, s, feats, n, W, sr, g, rgb);
    return hipGetLastError();
}

}  // namespace mrp
