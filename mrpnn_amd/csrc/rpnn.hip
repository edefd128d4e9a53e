// RPNN, the paper's comparison predictor: RenderCamera<true, RPNN> (core/volume.cu:198-228) ->
// NNPredict<RPNN> (core/render.cuh:808-875) -> RadiancePredict_RPNN
// (core/radiancePredict.cu:567-722) with the weights of core/NNWeight_RPNN.cuh
// (assets/rpnn_weights.bin, 1,296,401 floats).
//
// Per active pixel a descriptor of 10 levels x 225 density samples (5 x 5 x 9 grid in the light
// frame, spacing doubling per level) feeds a recurrent network: per level B = relu(V_l [226] +
// W_l C + b), C = relu(C + U_l B + b) with 200-wide states, then F0, F1 and F2; the radiance is
// max(exp(y) - 1, 0) for every channel.  1.29 M multiply-adds per pixel, 27x MRPNN's network.
//
// FP32 path (the reference's arithmetic): the march lists the active pixels; per chunk of them
// each level's 226 inputs are written to HBM and each layer is a GEMM over (pixels x 200
// outputs) whose every output sums its products from input 0 upwards, as Linear does, with the
// residual adds and ReLUs in the epilogues.  BF16 path: the descriptor as bf16 MFMA fragments
// (k_rpnn_desc), then one fused kernel runs the 31 layers on 16x16x32 MFMA (k_rpnn_mfma).
#include <algorithm>
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mrp_c_api.h"
#include "capi_internal.h"
#include "dev_common.hpp"
#include "fexp.h"

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return mrp_fail(MRP_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace mrp {

constexpr int kRL = 10, kRS = 225, kRIn = 226, kRB = 200;
constexpr int kRTotal = 1296401;
constexpr unsigned long long kRFnv = 0xad1ea655797fc91cull;

// element offsets of the asset (header order: V0..V9, W1..W9 with biases, U0..U9 with biases,
// F0, F1, F2 with biases)
struct RpnnLayout {
    size_t V[kRL], W[kRL], Wb[kRL], U[kRL], Ub[kRL], F0, F0b, F1, F1b, F2, F2b;
};
static RpnnLayout rpnn_layout() {
    RpnnLayout o{};
    size_t off = 0;
    for (int l = 0; l < kRL; ++l, off += (size_t)kRB * kRIn) o.V[l] = off;
    for (int l = 1; l < kRL; ++l) {
        o.W[l] = off;
        off += (size_t)kRB * kRB;
        o.Wb[l] = off;
        off += kRB;
    }
    for (int l = 0; l < kRL; ++l) {
        o.U[l] = off;
        off += (size_t)kRB * kRB;
        o.Ub[l] = off;
        off += kRB;
    }
    o.F0 = off;
    off += (size_t)kRB * kRB;
    o.F0b = off;
    off += kRB;
    o.F1 = off;
    off += (size_t)kRB * kRB;
    o.F1b = off;
    off += kRB;
    o.F2 = off;
    off += kRB;
    o.F2b = off;
    return o;
}

// One active pixel for the RPNN network: 32 bytes.
struct RpnnRec {
    int pixel;
    float pos[3];
    float X[3];
    float dist;
};

// RenderCamera<true, RPNN> up to the network: the same ray, march and basis as the MRPNN path
// (pixel_state); sky pixels are written at once, active ones listed (wave-aggregated).
// checker > 0 takes only the pixels with (row + col + checker - 1) even (the interactive
// renderer's checkerboard half); the others are left untouched.
__global__ __launch_bounds__(64) void k_rpnn_march(DevVolume vol, ViewDev vw, int checker, float4* __restrict__ out,
                                                   RpnnRec* __restrict__ recs, unsigned* __restrict__ count) {
    const int col = blockIdx.x * 8 + (threadIdx.x & 7);
    const int row = blockIdx.y * 8 + (threadIdx.x >> 3);
    const bool valid =
        row < vw.height && col < vw.width && (checker == 0 || ((row + col + checker - 1) & 1) == 0);
    PixelState st;
    st.active = 0;
    if (valid) st = pixel_state(vol, vw, row, col);
    const bool act = valid && st.active;
    const unsigned long long mask = __ballot(act);
    const int lane = threadIdx.x & 63;
    const size_t pix = (size_t)row * vw.width + col;
    if (valid && !act) {
        const v3 sky = sky_box(vol, normalize3(pixel_ray(vw, row, col, (uint32_t)pix)));
        out[pix] = make_float4(fmaxf(0.0f, sky.x), fmaxf(0.0f, sky.y), fmaxf(0.0f, sky.z), 0.0f);
    }
    if (!mask) return;
    const int lead = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if (lane == lead) base = atomicAdd(count, (unsigned)__popcll(mask));
    base = __shfl(base, lead);
    if (act) {
        RpnnRec r;
        r.pixel = (int)pix;
        r.pos[0] = st.pos.x;
        r.pos[1] = st.pos.y;
        r.pos[2] = st.pos.z;
        r.X[0] = st.X.x;
        r.X[1] = st.X.y;
        r.X[2] = st.X.z;
        r.dist = st.dist;
        recs[base + __popcll(mask & ((1ull << lane) - 1ull))] = r;
    }
}

// Sample k of level l of the descriptor (radiancePredict.cu:580-611): k = (z+2)*25 + (y+2)*5 + (x+2)
// on the 5 x 5 x 9 grid (x, y in [-2, 2], z in [-2, 6]), scaled by 2^(l-12), at mip
// max(min(l - 2, 9), 0); value density * alpha / 64.
__device__ inline float rpnn_sample(const DevVolume& vol, v3 pos, v3 eX, v3 eY, v3 LX, int l, int k, float alpha) {
    const int z = k / 25 - 2, y = (k / 5) % 5 - 2, x = k % 5 - 2;
    float scale = 0.5f * (0.5f / 1024.0f);
    for (int i = 0; i < l; ++i) scale *= 2.0f;
    const float cur = fmaxf(fminf((float)l - 2.0f, 9.0f), 0.0f);
    const int mip = (int)(cur + 0.001f);
    const v3 offset = (eX * (float)x + eY * (float)y + LX * (float)z) * scale;
    return mip_density(vol, mip, pos + offset) * alpha / 64.0f;
}

__device__ inline void rpnn_frame(v3 X, v3 LX, v3* eX, v3* eY) {
    *eX = normalize3(cross3(LX, X));
    *eY = normalize3(cross3(*eX, LX));
}

// level l's 226 inputs of rows [0, m): 225 samples, gamma = acos(dot(X, LX))
__global__ __launch_bounds__(256) void k_rpnn_input(DevVolume vol, const RpnnRec* __restrict__ recs, int m, int l,
                                                    v3 LX, float alpha, float* __restrict__ A) {
    const int p = blockIdx.x;
    if (p >= m) return;
    const RpnnRec r = recs[p];
    const v3 X = mk3(r.X[0], r.X[1], r.X[2]);
    v3 eX, eY;
    rpnn_frame(X, LX, &eX, &eY);
    const v3 pos = mk3(r.pos[0], r.pos[1], r.pos[2]);
    for (int k = threadIdx.x; k < kRIn; k += 256)
        A[(size_t)p * kRIn + k] = k < kRS ? rpnn_sample(vol, pos, eX, eY, LX, l, k, alpha) : acosf(dot3(X, LX));
}

// level l's 226 inputs from given rows (2250 + gamma)
__global__ __launch_bounds__(256) void k_rpnn_input_rows(const float* __restrict__ rows, int m, int l,
                                                         float* __restrict__ A) {
    const int p = blockIdx.x;
    if (p >= m) return;
    const float* row = rows + (size_t)p * (kRL * kRS + 1);
    for (int k = threadIdx.x; k < kRIn; k += 256) A[(size_t)p * kRIn + k] = k < kRS ? row[l * kRS + k] : row[kRL * kRS];
}

// Full descriptor rows (2250 + gamma) of all pixels of a view, zeros for sky (parity tooling)
__global__ __launch_bounds__(256) void k_rpnn_rows(DevVolume vol, const RpnnRec* __restrict__ recs, int m, v3 LX,
                                                   float alpha, float* __restrict__ rows) {
    const int p = blockIdx.x;
    if (p >= m) return;
    const RpnnRec r = recs[p];
    const v3 X = mk3(r.X[0], r.X[1], r.X[2]);
    v3 eX, eY;
    rpnn_frame(X, LX, &eX, &eY);
    const v3 pos = mk3(r.pos[0], r.pos[1], r.pos[2]);
    float* row = rows + (size_t)r.pixel * (kRL * kRS + 1);
    for (int k = threadIdx.x; k < kRL * kRS + 1; k += 256)
        row[k] = k < kRL * kRS ? rpnn_sample(vol, pos, eX, eY, LX, k / kRS, k % kRS, alpha) : acosf(dot3(X, LX));
}

// out[p][i] = sum_j A[p][j] W[i][j] (j ascending from 0), then the epilogue:
enum RpnnEpi {
    kEpiRelu = 0,        // relu(acc)                          V_0 A
    kEpiBias = 1,        // acc + b                            W_l C + b
    kEpiAddRelu = 2,     // relu(acc + add)                    V_l A + (W_l C + b)
    kEpiBiasRelu = 3,    // relu(acc + b)                      U_0 B + b, F0, F1
    kEpiResBiasRelu = 4, // relu(add + (acc + b)), in place    C + (U_l B + b)
    kEpiOut = 5          // max(exp(relu(acc + b)) - 1, 0)     F2 (exp_f32 of fexp.h)
};
constexpr int kGT = 64, kGK = 32;

template <int EPI>
__global__ __launch_bounds__(256) void k_rpnn_gemm(const float* __restrict__ A, int K, const float* __restrict__ W,
                                                   const float* __restrict__ bias, const float* add, float* out,
                                                   int m, int To) {
    __shared__ float sA[kGK][kGT + 1];
    __shared__ float sW[kGK][kGT + 1];
    const int p0 = blockIdx.x * kGT, o0 = blockIdx.y * kGT;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    float acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kGK) {
        for (int e = threadIdx.x; e < kGT * kGK; e += 256) {
            const int r = e / kGK, kk = e % kGK;
            const int k = k0 + kk;
            sA[kk][r] = (p0 + r < m && k < K) ? A[(size_t)(p0 + r) * K + k] : 0.0f;
            sW[kk][r] = (o0 + r < To && k < K) ? W[(size_t)(o0 + r) * K + k] : 0.0f;
        }
        __syncthreads();
        const int kn = min(kGK, K - k0);
        for (int kk = 0; kk < kn; ++kk) {
            float a[4], w[4];
            for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
            for (int j = 0; j < 4; ++j) w[j] = sW[kk][tx * 4 + j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + a[i] * w[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; ++i) {
        const int p = p0 + ty * 4 + i;
        if (p >= m) continue;
        for (int j = 0; j < 4; ++j) {
            const int o = o0 + tx * 4 + j;
            if (o >= To) continue;
            const size_t idx = (size_t)p * To + o;
            float v = acc[i][j];
            if (EPI == kEpiRelu) {
                v = v >= 0.0f ? v : 0.0f;
            } else if (EPI == kEpiBias) {
                v = v + bias[o];
            } else if (EPI == kEpiAddRelu) {
                v = v + add[idx];
                v = v >= 0.0f ? v : 0.0f;
            } else if (EPI == kEpiBiasRelu) {
                v = v + bias[o];
                v = v >= 0.0f ? v : 0.0f;
            } else if (EPI == kEpiResBiasRelu) {
                v = add[idx] + (v + bias[o]);
                v = v >= 0.0f ? v : 0.0f;
            } else {
                v = v + bias[o];
                v = v >= 0.0f ? v : 0.0f;
                const float e = exp_f32(v) - 1.0f;
                v = e > 0.0f ? e : 0.0f;
            }
            out[idx] = v;
        }
    }
}

// pixel = max(0, (y * light_color) * scatter_rate), distance (render.cuh:874, volume.cu:224)
__global__ __launch_bounds__(256) void k_rpnn_compose(const RpnnRec* __restrict__ recs, int m,
                                                      const float* __restrict__ y, v3 lc, v3 sr,
                                                      float4* __restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= m) return;
    const float v = y[p];
    const v3 r = (mk3(v, v, v) * lc) * sr;
    out[recs[p].pixel] = make_float4(fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z), fmaxf(0.0f, recs[p].dist));
}

// ------------------------------------------------------------------------------------------
// BF16 MFMA path: one wavefront = 16 pixels (lane = c + 16q), weights as A fragments (16
// output rows per M-tile, 13 tiles for the 200-wide states), activations as B fragments in the
// K-slot order where slot 8q+i <-> input 32k+4q+i and slot 8q+4+i <-> input 32k+16+4q+i, so a
// layer's D registers are the next layer's B operand.  A 512-thread workgroup (128 pixels)
// streams each layer's fragments through two LDS slots in K halves; the descriptor comes as
// ready-made B fragments from HBM (k_rpnn_desc).
// ------------------------------------------------------------------------------------------
typedef __bf16 rbf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 rbf16x2_t __attribute__((ext_vector_type(2)));
typedef float rf32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned ru32x4_t __attribute__((ext_vector_type(4)));

constexpr int kRNM = 13;                    // M-tiles of a 200-wide layer
constexpr int kRKV = 8, kRKS = 7;           // 32-wide K chunks of the level input (226) and a state (200)
constexpr int kRStages = kRL * 3 - 1 + 3;   // V, W (levels 1..9), U per level; F0, F1, F2

struct RpnnStage {
    int frag;  // first uint4 of the layer image
    int nk;    // K chunks
    int nm;    // M-tiles
    int bias;  // first float of the padded bias (16 * nm), -1 without
};
struct RpnnStages {
    RpnnStage s[kRStages];
};

__device__ inline uint32_t rpack_bf16x2(float a, float b) {  // round to nearest even
    const rbf16x2_t v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ inline ru32x4_t rpack8(const float* v) {
    return ru32x4_t{rpack_bf16x2(v[0], v[1]), rpack_bf16x2(v[2], v[3]), rpack_bf16x2(v[4], v[5]),
                    rpack_bf16x2(v[6], v[7])};
}
__device__ inline rf32x4_t rmfma(ru32x4_t a, ru32x4_t b, rf32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(rbf16x8_t, a), __builtin_bit_cast(rbf16x8_t, b),
                                                   c, 0, 0, 0);
}
// B fragments of a 200-wide state held in D layout (13 tiles): chunk k = tiles 2k, 2k+1
__device__ inline void rstate_frags(const rf32x4_t (&d)[kRNM], ru32x4_t (&f)[kRKS]) {
    for (int k = 0; k < kRKS; ++k) {
        const rf32x4_t a = d[2 * k];
        const rf32x4_t b = 2 * k + 1 < kRNM ? d[2 * k + 1] : rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        f[k] = ru32x4_t{rpack_bf16x2(a[0], a[1]), rpack_bf16x2(a[2], a[3]), rpack_bf16x2(b[0], b[1]),
                        rpack_bf16x2(b[2], b[3])};
    }
}
// Weight staging: each layer is split in K halves ([0, 4) and [4, nk) chunks) and a half's
// fragments go from HBM / L2 straight into one of two LDS slots (global_load_lds, no registers),
// issued one half ahead: while the workgroup multiplies with the half in one slot, the next
// half streams into the other.  A slot holds 13 M-tiles x 4 K chunks (52 KB); fragment
// (t, k) of a half sits at (t * nkh + k - K0) * 64 + lane.
typedef __attribute__((address_space(3))) void* lds_void_ptr;
__device__ inline void rdma_half(ru32x4_t* slot, const ru32x4_t* __restrict__ img, const RpnnStage& st, int K0,
                                 int K1, int wave) {
    const int nkh = K1 - K0, lane = threadIdx.x & 63;
    for (int b = wave; b < st.nm * nkh; b += 8) {  // one 1 KB block per wavefront instruction
        const int t = b / nkh, kk = b - t * nkh;
        __builtin_amdgcn_global_load_lds((const void*)(img + st.frag + (t * st.nk + K0 + kk) * 64 + lane),
                                         (lds_void_ptr)(slot + b * 64), 16, 0, 0);
    }
}
// the end of a half: this wavefront's copies have landed, every wavefront is done with the slot
// it read, and the other slot's fragments are visible to all
__device__ inline void rhalf_end() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}
// acc[t] += W[t][K0..K1) x in[K0..K1) for the 13 M-tiles from a slot.  Each A fragment
// register is refilled with the next K chunk's fragment right after its MFMA (scheduling
// group barriers pin that interleave), so the LDS latency hides behind the MFMAs.
template <int K0, int K1, int NK>
__device__ inline void rcompute(const ru32x4_t* slot, const ru32x4_t (&in)[NK], rf32x4_t (&acc)[kRNM], int lane) {
    constexpr int nkh = K1 - K0;
    ru32x4_t a[kRNM];
#pragma unroll
    for (int t = 0; t < kRNM; ++t) a[t] = slot[t * nkh * 64 + lane];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = K0; k < K1; ++k) {
#pragma unroll
        for (int t = 0; t < kRNM; ++t) {
            acc[t] = rmfma(a[t], in[k], acc[t]);
            if (k + 1 < K1) a[t] = slot[(t * nkh + k + 1 - K0) * 64 + lane];
        }
#pragma unroll
        for (int t = 0; t < kRNM; ++t) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                  // one MFMA
            if (k + 1 < K1) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}
__device__ inline void rbias_relu(rf32x4_t (&acc)[kRNM], const float* __restrict__ bias, int q, bool relu) {
    for (int t = 0; t < kRNM; ++t) {
        const float4 b = bias ? *(const float4*)(bias + 16 * t + 4 * q) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rf32x4_t v = acc[t] + rf32x4_t{b.x, b.y, b.z, b.w};
        if (relu)
            for (int i = 0; i < 4; ++i) v[i] = v[i] >= 0.0f ? v[i] : 0.0f;
        acc[t] = v;
    }
}

// The descriptor as the first layer's B fragments, bf16, in HBM: item ((t * 10 + l) * 8 + k) * 64
// + lane holds the 8 inputs of K chunk k of level l for pixel 16t + (lane & 15), slots of
// q = lane >> 4.  A separate, fully occupied kernel hides the latency of the 2250 trilinear
// samples per pixel that the MFMA kernel would otherwise wait for.
//
// One workgroup per (16-pixel tile, level).  A level's 225 samples form a lattice of spacing
// 1/16, 1/8, then 1/4 texel of its mip (levels 0, 1, then 2..9 on mip l-2), so they reach at
// most 5 texels per axis: each pixel's 5 x 5 x 5 box of the mip (zeros outside the grid,
// pre-multiplied by alpha / 64) is staged in LDS once, then every sample is an LDS trilinear
// at texel coordinates u0 + x a + y b + z c (fused multiply-adds; the values differ from the
// FP32 path's by float rounding, far below the bf16 step they are stored at).
constexpr int kRBox = 125;       // 5 x 5 x 5 texels
constexpr int kRDescStride = 260;  // floats per pixel of the staged level input (256 slots + pad)
__global__ __launch_bounds__(256) void k_rpnn_desc(DevVolume vol, const RpnnRec* __restrict__ recs, int m, v3 LX,
                                                   float alpha, ru32x4_t* __restrict__ desc) {
    __shared__ float sBox[16 * kRBox];
    __shared__ __align__(16) float sIn[16 * kRDescStride];
    __shared__ float sPix[16][16];  // u0, a, b, c (texel units), box origin, gamma
    const int t = blockIdx.x / kRL, l = blockIdx.x % kRL;
    const int mip = l < 2 ? 0 : l - 2;
    const int res = 256 >> mip;
    const float* __restrict__ grid = vol.mips[mip];
    if (threadIdx.x < 16) {
        const int p = 16 * t + threadIdx.x;
        float* P = sPix[threadIdx.x];
        if (p < m) {
            const RpnnRec r = recs[p];
            const v3 X = mk3(r.X[0], r.X[1], r.X[2]);
            v3 eX, eY;
            rpnn_frame(X, LX, &eX, &eY);
            const float sp = ldexpf(1.0f, l - 12) * (float)res;  // lattice spacing in texels
            const v3 u0 = (mk3(r.pos[0], r.pos[1], r.pos[2]) + 0.5f) * (float)res - 0.5f;
            const v3 a = eX * sp, b = eY * sp, c = LX * sp;
            const float u[3] = {u0.x, u0.y, u0.z}, A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, C[3] = {c.x, c.y, c.z};
            for (int i = 0; i < 3; ++i) {
                // lowest coordinate of the lattice on this axis (x, y in [-2, 2], z in [-2, 6])
                const float lo = floorf(u[i] - 2.0f * fabsf(A[i]) - 2.0f * fabsf(B[i]) + fminf(-2.0f * C[i], 6.0f * C[i]));
                P[i] = u[i] - lo;  // lattice origin relative to the box
                P[3 + i] = A[i];
                P[6 + i] = B[i];
                P[9 + i] = C[i];
                P[12 + i] = lo;
            }
            P[15] = acosf(dot3(X, LX));
        } else {
            for (int i = 0; i < 16; ++i) P[i] = 0.0f;
            P[12] = P[13] = P[14] = -1.0e6f;  // box entirely outside the grid: zeros
        }
    }
    __syncthreads();
    // the boxes: 16 pixels x 25 (x, y) rows of 5 texels along z, loads at clamped in-grid
    // indices, zero outside the grid
    const float scale = alpha / 64.0f;
    for (int i = threadIdx.x; i < 16 * 25; i += 256) {
        const int px = i / 25, rr = i - px * 25;
        const int x = (int)sPix[px][12] + rr / 5, y = (int)sPix[px][13] + rr % 5, z0 = (int)sPix[px][14];
        const bool row_in = x >= 0 && y >= 0 && x < res && y < res;
        const float* __restrict__ src = grid + ((size_t)(row_in ? x : 0) * res + (row_in ? y : 0)) * res;
        float* dst = sBox + px * kRBox + rr * 5;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int z = z0 + k;
            const float v = src[min(max(z, 0), res - 1)] * scale;
            dst[k] = row_in && z >= 0 && z < res ? v : 0.0f;
        }
    }
    // slots 225 (gamma) .. 255 of the level input
    for (int i = threadIdx.x; i < 16 * 31; i += 256) {
        const int px = i / 31, s = kRS + i % 31;
        sIn[px * kRDescStride + s] = s == kRS ? sPix[px][15] : 0.0f;
    }
    __syncthreads();
    // 16 pixels x 45 lattice rows (y, z) of 5 samples along x; lanes take different pixels
    for (int i = threadIdx.x; i < 16 * 45; i += 256) {
        const int px = i & 15, row = i >> 4;
        const float* P = sPix[px];
        const float fy = (float)(row % 5 - 2), fz = (float)(row / 5 - 2);
        float base[3];
        for (int a = 0; a < 3; ++a) base[a] = fmaf(fz, P[9 + a], fmaf(fy, P[6 + a], P[a]));
        const float* box = sBox + px * kRBox;
        float* dst = sIn + px * kRDescStride + row * 5;
#pragma unroll
        for (int xi = 0; xi < 5; ++xi) {
            const float fx = (float)(xi - 2);
            float w[3];
            int o[3];
            for (int a = 0; a < 3; ++a) {
                // box coordinate; the clamp only absorbs rounding at the box's lower face (the
                // lattice spans < 3.5 texels per axis from its lowest coordinate)
                const float d = __builtin_amdgcn_fmed3f(fmaf(fx, P[3 + a], base[a]), 0.0f, 3.99f);
                w[a] = __builtin_amdgcn_fractf(d);
                o[a] = (int)d;
            }
            const float* b0 = box + (__mul24(o[0], 25) + __mul24(o[1], 5) + o[2]);  // o in [0, 3]
            const float c00 = fmaf(w[2], b0[1] - b0[0], b0[0]), c01 = fmaf(w[2], b0[6] - b0[5], b0[5]);
            const float c10 = fmaf(w[2], b0[26] - b0[25], b0[25]), c11 = fmaf(w[2], b0[31] - b0[30], b0[30]);
            const float c0 = fmaf(w[1], c01 - c00, c00), c1 = fmaf(w[1], c11 - c10, c10);
            dst[xi] = fmaf(w[0], c1 - c0, c0);
        }
    }
    __syncthreads();
    // two fragments per thread: item k * 64 + lane of this (tile, level)
    for (int j = threadIdx.x; j < kRKV * 64; j += 256) {
        const int k = j >> 6, lane = j & 63, c = lane & 15, q = lane >> 4;
        const float* src = sIn + c * kRDescStride + 32 * k + 4 * q;
        const float4 lo4 = *(const float4*)src, hi4 = *(const float4*)(src + 16);
        const float v[8] = {lo4.x, lo4.y, lo4.z, lo4.w, hi4.x, hi4.y, hi4.z, hi4.w};
        // streamed once: kept out of the caches the network kernel's weights live in
        __builtin_nontemporal_store(16 * t + c < m ? rpack8(v) : ru32x4_t{0u, 0u, 0u, 0u},
                                    desc + (size_t)blockIdx.x * (kRKV * 64) + j);
    }
}

// The same fragments from given rows (2250 descriptor values, gamma): the BF16 network alone
__global__ __launch_bounds__(256) void k_rpnn_desc_rows(const float* __restrict__ rows, int m,
                                                        ru32x4_t* __restrict__ desc) {
    const long long item = (long long)blockIdx.x * 256 + threadIdx.x;
    const int tiles = (m + 15) / 16;
    if (item >= (long long)tiles * kRL * kRKV * 64) return;
    const int lane = (int)(item & 63), k = (int)((item >> 6) % kRKV), l = (int)((item / (64 * kRKV)) % kRL);
    const int t = (int)(item / (64 * kRKV * kRL));
    const int p = 16 * t + (lane & 15), q = lane >> 4;
    float v[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (p < m) {
        const float* row = rows + (size_t)p * (kRL * kRS + 1);
        for (int j = 0; j < 8; ++j) {
            const int idx = 32 * k + (j < 4 ? 4 * q + j : 16 + 4 * q + j - 4);
            v[j] = idx < kRS ? row[l * kRS + idx] : (idx == kRS ? row[kRL * kRS] : 0.0f);
        }
    }
    desc[item] = rpack8(v);
}

// Stage indices in the image (V0, U0, then V_l, W_l, U_l per level, F0, F1, F2)
__device__ constexpr int rs_v(int l) { return l == 0 ? 0 : 3 * l - 1; }
__device__ constexpr int rs_w(int l) { return 3 * l; }
__device__ constexpr int rs_u(int l) { return l == 0 ? 1 : 3 * l + 1; }
constexpr int kRsF0 = 3 * kRL - 1;

// The fused network: a 128-pixel workgroup of 8 wavefronts with 16 pixels each, the 200-wide
// states in registers (D layout), 64 weight halves per tile streamed through two LDS slots.
__global__ __launch_bounds__(512) void k_rpnn_mfma(const RpnnRec* __restrict__ recs, int m,
                                                   const ru32x4_t* __restrict__ desc,
                                                   const ru32x4_t* __restrict__ img, const float* __restrict__ bias,
                                                   const RpnnStage* __restrict__ stages, v3 lc, v3 sr,
                                                   float4* __restrict__ out, float* __restrict__ y_out) {
    __shared__ ru32x4_t sA[kRNM * 4 * 64];
    __shared__ ru32x4_t sB[kRNM * 4 * 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, q = lane >> 4;
    const int tiles = (m + 127) / 128;
    if ((int)blockIdx.x >= tiles) return;
    rdma_half(sA, img, stages[rs_v(0)], 0, 4, wave);
    rhalf_end();
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int t16 = tile * 8 + wave;  // this wave's 16 pixels
        const ru32x4_t* dsc = desc + (size_t)t16 * kRL * kRKV * 64 + lane;
        rf32x4_t C[kRNM], acc[kRNM];
        for (int l = 0; l < kRL; ++l) {
            // B = relu(V_l [desc_l, gamma] (+ W_l C + b)); the level's descriptor fragments are
            // requested first (read once: non-temporal, so they do not evict weights from L2)
            ru32x4_t dfr[kRKV];
            for (int k = 0; k < kRKV; ++k)
                dfr[k] = (16 * t16 < m) ? __builtin_nontemporal_load(dsc + (l * kRKV + k) * 64)
                                        : ru32x4_t{0u, 0u, 0u, 0u};
            const RpnnStage& sv = stages[rs_v(l)];
            rdma_half(sB, img, sv, 4, kRKV, wave);
            for (int t = 0; t < kRNM; ++t) acc[t] = rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
            rcompute<0, 4, kRKV>(sA, dfr, acc, lane);
            rhalf_end();
            if (l > 0) {
                const RpnnStage& sw = stages[rs_w(l)];
                rdma_half(sA, img, sw, 0, 4, wave);
                rcompute<4, kRKV, kRKV>(sB, dfr, acc, lane);
                rhalf_end();
                ru32x4_t Cf[kRKS];
                rstate_frags(C, Cf);
                rdma_half(sB, img, sw, 4, kRKS, wave);
                rcompute<0, 4, kRKS>(sA, Cf, acc, lane);
                rhalf_end();
                rdma_half(sA, img, stages[rs_u(l)], 0, 4, wave);
                rcompute<4, kRKS, kRKS>(sB, Cf, acc, lane);
                rbias_relu(acc, bias + sw.bias, q, true);
            } else {
                rdma_half(sA, img, stages[rs_u(l)], 0, 4, wave);
                rcompute<4, kRKV, kRKV>(sB, dfr, acc, lane);
                rbias_relu(acc, nullptr, q, true);
            }
            rhalf_end();
            // C = relu((C +) U_l B + b), accumulated in place
            const RpnnStage& su = stages[rs_u(l)];
            ru32x4_t Bf[kRKS];
            rstate_frags(acc, Bf);
            if (l == 0)
                for (int t = 0; t < kRNM; ++t) C[t] = rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
            rdma_half(sB, img, su, 4, kRKS, wave);
            rcompute<0, 4, kRKS>(sA, Bf, C, lane);
            rhalf_end();
            rdma_half(sA, img, stages[l + 1 < kRL ? rs_v(l + 1) : kRsF0], 0, 4, wave);
            rcompute<4, kRKS, kRKS>(sB, Bf, C, lane);
            rbias_relu(C, bias + su.bias, q, true);
            rhalf_end();
        }
        // F0, F1 (ReLU), F2 (one output row), radiance max(exp(y) - 1, 0)
        {
            const RpnnStage& s0 = stages[kRsF0];
            ru32x4_t Cf[kRKS];
            rstate_frags(C, Cf);
            for (int t = 0; t < kRNM; ++t) acc[t] = rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
            rdma_half(sB, img, s0, 4, kRKS, wave);
            rcompute<0, 4, kRKS>(sA, Cf, acc, lane);
            rhalf_end();
            rdma_half(sA, img, stages[kRsF0 + 1], 0, 4, wave);
            rcompute<4, kRKS, kRKS>(sB, Cf, acc, lane);
            rbias_relu(acc, bias + s0.bias, q, true);
            rhalf_end();
        }
        {
            const RpnnStage& s1 = stages[kRsF0 + 1];
            ru32x4_t Bf[kRKS];
            rstate_frags(acc, Bf);
            for (int t = 0; t < kRNM; ++t) C[t] = rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
            rdma_half(sB, img, s1, 4, kRKS, wave);
            rcompute<0, 4, kRKS>(sA, Bf, C, lane);
            rhalf_end();
            rdma_half(sA, img, stages[kRsF0 + 2], 0, 4, wave);
            rcompute<4, kRKS, kRKS>(sB, Bf, C, lane);
            rbias_relu(C, bias + s1.bias, q, true);
            rhalf_end();
        }
        const RpnnStage& s2 = stages[kRsF0 + 2];
        ru32x4_t Cf[kRKS];
        rstate_frags(C, Cf);
        rdma_half(sB, img, s2, 4, kRKS, wave);
        rf32x4_t y = rf32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 4; ++k) y = rmfma(sA[k * 64 + lane], Cf[k], y);
        rhalf_end();
        if (tile + (int)gridDim.x < tiles) rdma_half(sA, img, stages[rs_v(0)], 0, 4, wave);  // the next tile
        for (int k = 4; k < kRKS; ++k) y = rmfma(sB[(k - 4) * 64 + lane], Cf[k], y);
        const int p = 16 * t16 + c;
        if (p < m && q == 0) {
            float v = y[0] + bias[s2.bias];
            v = v >= 0.0f ? v : 0.0f;
            float o = __expf(v) - 1.0f;
            o = o > 0.0f ? o : 0.0f;
            if (y_out) {
                y_out[p] = o;
            } else {
                const v3 rr = (mk3(o, o, o) * lc) * sr;
                const RpnnRec& r = recs[p];
                out[r.pixel] = make_float4(fmaxf(0.0f, rr.x), fmaxf(0.0f, rr.y), fmaxf(0.0f, rr.z), fmaxf(0.0f, r.dist));
            }
        }
        rhalf_end();
    }
}

}  // namespace mrp

using namespace mrp;

namespace mrp {
uint16_t host_f32_to_bf16(float f);
}

// Fragment images and padded biases of the MFMA path, in stage order (rpnn_stage_list).
static void rpnn_pack(const std::vector<float>& w, const RpnnLayout& L, std::vector<ru32x4_t>& img,
                      std::vector<float>& bias, RpnnStages& st) {
    int si = 0;
    auto add = [&](size_t wo, long bo, int M, int K) {
        const int nm = (M + 15) / 16, nk = (K + 31) / 32;
        RpnnStage& s = st.s[si++];
        s.frag = (int)img.size();
        s.nk = nk;
        s.nm = nm;
        s.bias = bo < 0 ? -1 : (int)bias.size();
        for (int mt = 0; mt < nm; ++mt)
            for (int kc = 0; kc < nk; ++kc)
                for (int lane = 0; lane < 64; ++lane) {
                    const int qq = lane >> 4, row = 16 * mt + (lane & 15);
                    uint16_t h[8];
                    for (int slot = 0; slot < 8; ++slot) {
                        const int k = 32 * kc + (slot < 4 ? 4 * qq + slot : 16 + 4 * qq + slot - 4);
                        h[slot] = host_f32_to_bf16(row < M && k < K ? w[wo + (size_t)row * K + k] : 0.0f);
                    }
                    img.push_back(ru32x4_t{(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                                           (uint32_t)h[4] | ((uint32_t)h[5] << 16), (uint32_t)h[6] | ((uint32_t)h[7] << 16)});
                }
        if (bo >= 0)
            for (int r = 0; r < 16 * nm; ++r) bias.push_back(r < M ? w[(size_t)bo + r] : 0.0f);
    };
    for (int l = 0; l < kRL; ++l) {
        add(L.V[l], -1, kRB, kRIn);
        if (l > 0) add(L.W[l], (long)L.Wb[l], kRB, kRB);
        add(L.U[l], (long)L.Ub[l], kRB, kRB);
    }
    add(L.F0, (long)L.F0b, kRB, kRB);
    add(L.F1, (long)L.F1b, kRB, kRB);
    add(L.F2, (long)L.F2b, 1, kRB);
}

struct RpnnState {
    float* d_w = nullptr;  // fp32 asset, header order
    RpnnLayout lay;
    ru32x4_t* d_img = nullptr;  // bf16 fragments of the MFMA path
    float* d_bias = nullptr;
    RpnnStages stages;
    RpnnStage* d_stages = nullptr;
    ru32x4_t* d_desc = nullptr;  // descriptor fragments of one chunk of pixels
    int desc_cap = 0;            // pixels
    size_t rec_cap = 0;
    RpnnRec* d_recs = nullptr;
    unsigned* d_count = nullptr;
    int chunk = 0;
    float *d_A = nullptr, *d_B = nullptr, *d_C = nullptr, *d_T = nullptr, *d_y = nullptr;
};

void mrp_rpnn_free(mrp_ctx* c) {
    RpnnState* s = c->rpnn;
    if (!s) return;
    (void)hipFree(s->d_w);
    (void)hipFree(s->d_img);
    (void)hipFree(s->d_bias);
    (void)hipFree(s->d_desc);
    (void)hipFree(s->d_stages);
    (void)hipFree(s->d_recs);
    (void)hipFree(s->d_count);
    (void)hipFree(s->d_A);
    (void)hipFree(s->d_B);
    (void)hipFree(s->d_C);
    (void)hipFree(s->d_T);
    (void)hipFree(s->d_y);
    delete s;
    c->rpnn = nullptr;
}

static std::string default_rpnn_path() {
    Dl_info info;
    if (dladdr((void*)&mrp_ctx_load_rpnn, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        return (s == std::string::npos ? std::string(".") : p.substr(0, s)) + "/assets/rpnn_weights.bin";
    }
    return "assets/rpnn_weights.bin";
}

extern "C" int mrp_ctx_load_rpnn(mrp_ctx* c, const char* path) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (c->rpnn) return MRP_OK;
    const std::string p = path ? path : default_rpnn_path();
    FILE* f = fopen(p.c_str(), "rb");
    if (!f) return mrp_fail(MRP_ERR_IO, "cannot open RPNN weight asset %s", p.c_str());
    std::vector<float> w(kRTotal);
    size_t got = fread(w.data(), sizeof(float), w.size(), f);
    char extra;
    size_t more = fread(&extra, 1, 1, f);
    fclose(f);
    if (got != w.size() || more) return mrp_fail(MRP_ERR_IO, "RPNN weight asset %s: expected %d floats", p.c_str(), kRTotal);
    uint64_t h = 0xcbf29ce484222325ull;
    const unsigned char* b = (const unsigned char*)w.data();
    for (size_t i = 0; i < w.size() * 4; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    if (h != kRFnv) return mrp_fail(MRP_ERR_IO, "RPNN weight asset %s: checksum mismatch", p.c_str());
    RpnnState* s = new RpnnState();
    s->lay = rpnn_layout();
    std::vector<ru32x4_t> img;
    std::vector<float> bias;
    rpnn_pack(w, s->lay, img, bias, s->stages);
    if (hipMalloc(&s->d_w, w.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(s->d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->d_img, img.size() * sizeof(ru32x4_t)) != hipSuccess ||
        hipMemcpy(s->d_img, img.data(), img.size() * sizeof(ru32x4_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->d_bias, bias.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(s->d_bias, bias.data(), bias.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->d_stages, sizeof(RpnnStages)) != hipSuccess ||
        hipMemcpy(s->d_stages, &s->stages, sizeof(RpnnStages), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&s->d_count, sizeof(unsigned)) != hipSuccess) {
        c->rpnn = s;
        mrp_rpnn_free(c);
        return mrp_fail(MRP_ERR_NOMEM, "RPNN weights upload failed");
    }
    c->rpnn = s;
    return MRP_OK;
}

// active-pixel list of a view (march) and the sky pixels of out4; returns the count on the host
static int rpnn_march(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int checker, float* dev_out4, int* n_active) {
    RpnnState* s = c->rpnn;
    const size_t n = (size_t)vw->width * vw->height;
    if (n > s->rec_cap) {
        (void)hipFree(s->d_recs);
        s->d_recs = nullptr;
        if (hipMalloc(&s->d_recs, n * sizeof(RpnnRec)) != hipSuccess) {
            s->rec_cap = 0;
            return mrp_fail(MRP_ERR_NOMEM, "RPNN record list");
        }
        s->rec_cap = n;
    }
    HIP_TRY(hipMemsetAsync(s->d_count, 0, sizeof(unsigned), c->stream));
    dim3 grid((vw->width + 7) / 8, (vw->height + 7) / 8);
    hipLaunchKernelGGL(k_rpnn_march, grid, dim3(64), 0, c->stream, mrp_dev_volume(v), mrp_to_dev(vw), checker,
                       (float4*)dev_out4, s->d_recs, s->d_count);
    HIP_TRY(hipGetLastError());
    unsigned cnt = 0;
    HIP_TRY(hipMemcpyAsync(&cnt, s->d_count, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *n_active = (int)cnt;
    return MRP_OK;
}

static int rpnn_buffers(RpnnState* s, int chunk) {
    if (s->chunk >= chunk) return MRP_OK;
    (void)hipFree(s->d_A);
    (void)hipFree(s->d_B);
    (void)hipFree(s->d_C);
    (void)hipFree(s->d_T);
    (void)hipFree(s->d_y);
    s->d_A = s->d_B = s->d_C = s->d_T = s->d_y = nullptr;
    s->chunk = 0;
    if (hipMalloc(&s->d_A, (size_t)chunk * kRIn * 4) != hipSuccess ||
        hipMalloc(&s->d_B, (size_t)chunk * kRB * 4) != hipSuccess ||
        hipMalloc(&s->d_C, (size_t)chunk * kRB * 4) != hipSuccess ||
        hipMalloc(&s->d_T, (size_t)chunk * kRB * 4) != hipSuccess || hipMalloc(&s->d_y, (size_t)chunk * 4) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "RPNN activations");
    s->chunk = chunk;
    return MRP_OK;
}

template <int EPI>
static hipError_t gemm(hipStream_t st, const float* A, int K, const float* W, const float* bias, const float* add,
                       float* out, int m, int To) {
    dim3 grid((m + kGT - 1) / kGT, (To + kGT - 1) / kGT);
    hipLaunchKernelGGL(k_rpnn_gemm<EPI>, grid, dim3(256), 0, st, A, K, W, bias, add, out, m, To);
    return hipGetLastError();
}

// The FP32 network over m inputs in chunks: from the march's records (view set) or from given
// rows (rows set); the outputs are composed into out4 (records) or copied to y_out (rows).
static int rpnn_fp32(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int m, float* dev_out4, const float* rows,
                     float* y_out) {
    RpnnState* s = c->rpnn;
    const int chunk = 65536;
    int rc = rpnn_buffers(s, chunk);
    if (rc) return rc;
    DevVolume dv{};
    v3 LX = mk3(0, 0, 1), lc = mk3(1, 1, 1), sr = mk3(1, 1, 1);
    if (vw) {
        dv = mrp_dev_volume(v);
        LX = normalize3(mk3(vw->light_dir[0], vw->light_dir[1], vw->light_dir[2]));
        lc = mk3(vw->light_color[0], vw->light_color[1], vw->light_color[2]);
        sr = mk3(vw->scatter_rate[0], vw->scatter_rate[1], vw->scatter_rate[2]);
    }
    const float* w = s->d_w;
    const RpnnLayout& L = s->lay;
    hipStream_t st = c->stream;
    for (int off = 0; off < m; off += chunk) {
        const int n = m - off < chunk ? m - off : chunk;
        const RpnnRec* recs = s->d_recs + off;
        for (int l = 0; l < kRL; ++l) {
            if (rows)
                hipLaunchKernelGGL(k_rpnn_input_rows, dim3(n), dim3(256), 0, st, rows + (size_t)off * (kRL * kRS + 1),
                                   n, l, s->d_A);
            else
                hipLaunchKernelGGL(k_rpnn_input, dim3(n), dim3(256), 0, st, dv, recs, n, l, LX, vw->alpha, s->d_A);
            HIP_TRY(hipGetLastError());
            if (l == 0) {
                HIP_TRY(gemm<kEpiRelu>(st, s->d_A, kRIn, w + L.V[0], nullptr, nullptr, s->d_B, n, kRB));
                HIP_TRY(gemm<kEpiBiasRelu>(st, s->d_B, kRB, w + L.U[0], w + L.Ub[0], nullptr, s->d_C, n, kRB));
            } else {
                HIP_TRY(gemm<kEpiBias>(st, s->d_C, kRB, w + L.W[l], w + L.Wb[l], nullptr, s->d_T, n, kRB));
                HIP_TRY(gemm<kEpiAddRelu>(st, s->d_A, kRIn, w + L.V[l], nullptr, s->d_T, s->d_B, n, kRB));
                HIP_TRY(gemm<kEpiResBiasRelu>(st, s->d_B, kRB, w + L.U[l], w + L.Ub[l], s->d_C, s->d_C, n, kRB));
            }
        }
        HIP_TRY(gemm<kEpiBiasRelu>(st, s->d_C, kRB, w + L.F0, w + L.F0b, nullptr, s->d_T, n, kRB));
        HIP_TRY(gemm<kEpiBiasRelu>(st, s->d_T, kRB, w + L.F1, w + L.F1b, nullptr, s->d_B, n, kRB));
        HIP_TRY(gemm<kEpiOut>(st, s->d_B, kRB, w + L.F2, w + L.F2b, nullptr, rows ? y_out + off : s->d_y, n, 1));
        if (!rows) {
            hipLaunchKernelGGL(k_rpnn_compose, dim3((n + 255) / 256), dim3(256), 0, st, recs, n, s->d_y, lc, sr,
                               (float4*)dev_out4);
            HIP_TRY(hipGetLastError());
        }
    }
    return MRP_OK;
}

// The BF16 network over m inputs in chunks: descriptor fragments from the march's records (view
// set) or from given rows (rows set), then the fused MFMA network; outputs composed into out4
// (records) or written to y_out (rows).
static int rpnn_bf16(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int m, float* dev_out4, const float* rows,
                     float* y_out) {
    RpnnState* s = c->rpnn;
    v3 LX = mk3(0, 0, 1), lc = mk3(1, 1, 1), sr = mk3(1, 1, 1);
    DevVolume dv{};
    if (vw) {
        dv = mrp_dev_volume(v);
        LX = normalize3(mk3(vw->light_dir[0], vw->light_dir[1], vw->light_dir[2]));
        lc = mk3(vw->light_color[0], vw->light_color[1], vw->light_color[2]);
        sr = mk3(vw->scatter_rate[0], vw->scatter_rate[1], vw->scatter_rate[2]);
    }
    // descriptor fragments of c->rpnn_chunk pixels at a time (5 KB per pixel, bf16); the buffer
    // grows geometrically (at least doubling, at most one pass) so that active counts creeping
    // up frame after frame reallocate a few times only
    const int chunk = c->rpnn_chunk;
    const int fit = m < chunk ? (m + 127) / 128 * 128 : chunk;
    if (s->desc_cap < fit) {
        const int need = std::min(chunk, std::max(fit, 2 * s->desc_cap));
        (void)hipFree(s->d_desc);
        s->d_desc = nullptr;
        s->desc_cap = 0;
        if (hipMalloc(&s->d_desc, (size_t)need / 16 * kRL * kRKV * 64 * sizeof(ru32x4_t)) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "RPNN descriptor buffer");
        s->desc_cap = need;
    }
    const int cus = c->props.multiProcessorCount > 0 ? c->props.multiProcessorCount : 256;
    for (int off = 0; off < m; off += chunk) {
        const int n = m - off < chunk ? m - off : chunk;
        if (rows) {
            const long long items = (long long)((n + 15) / 16) * kRL * kRKV * 64;
            hipLaunchKernelGGL(k_rpnn_desc_rows, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, c->stream,
                               rows + (size_t)off * (kRL * kRS + 1), n, s->d_desc);
        } else {
            hipLaunchKernelGGL(k_rpnn_desc, dim3((unsigned)((n + 15) / 16 * kRL)), dim3(256), 0, c->stream, dv,
                               s->d_recs + off, n, LX, vw->alpha, s->d_desc);
        }
        HIP_TRY(hipGetLastError());
        const int tiles = (n + 127) / 128;
        hipLaunchKernelGGL(k_rpnn_mfma, dim3(tiles < cus ? tiles : cus), dim3(512), 0, c->stream,
                           rows ? nullptr : s->d_recs + off, n, s->d_desc, s->d_img, s->d_bias, s->d_stages, lc, sr,
                           (float4*)dev_out4, rows ? y_out + off : nullptr);
        HIP_TRY(hipGetLastError());
    }
    return MRP_OK;
}

extern "C" int mrp_render_rpnn(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int precision, float* dev_out4) {
    return mrp_render_rpnn_checker(c, v, vw, precision, dev_out4, 0);
}

int mrp_render_rpnn_checker(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int precision, float* dev_out4,
                            int checker) {
    if (!c || !v || !vw || !dev_out4) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (checker < 0 || checker > 2) return mrp_fail(MRP_ERR_INVALID, "checker must be 0, 1 or 2");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (precision != MRP_FP32 && precision != MRP_BF16)
        return mrp_fail(MRP_ERR_INVALID, "RPNN precision %d not available (MRP_FP32, MRP_BF16)", precision);
    if (!c->rpnn && (rc = mrp_ctx_load_rpnn(c, nullptr))) return rc;
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before rendering");
    int m = 0;
    rc = rpnn_march(c, v, vw, checker, dev_out4, &m);
    if (rc || m == 0) return rc;
    if (precision == MRP_FP32) return rpnn_fp32(c, v, vw, m, dev_out4, nullptr, nullptr);
    return rpnn_bf16(c, v, vw, m, dev_out4, nullptr, nullptr);
}

extern "C" int mrp_rpnn_predict_rows(mrp_ctx* c, const float* dev_rows, int n, int precision, float* dev_y) {
    if (!c || !dev_rows || !dev_y || n < 0) return mrp_fail(MRP_ERR_INVALID, "bad arguments");
    if (precision != MRP_FP32 && precision != MRP_BF16)
        return mrp_fail(MRP_ERR_INVALID, "RPNN precision %d not available (MRP_FP32, MRP_BF16)", precision);
    int rc;
    if (!c->rpnn && (rc = mrp_ctx_load_rpnn(c, nullptr))) return rc;
    if (n == 0) return MRP_OK;
    if (precision == MRP_BF16) return rpnn_bf16(c, nullptr, nullptr, n, nullptr, dev_rows, dev_y);
    return rpnn_fp32(c, nullptr, nullptr, n, nullptr, dev_rows, dev_y);
}

extern "C" int mrp_rpnn_features(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, float* dev_rows, float* dev_out4) {
    if (!c || !v || !vw || !dev_rows || !dev_out4) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (!c->rpnn && (rc = mrp_ctx_load_rpnn(c, nullptr))) return rc;
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before rendering");
    const size_t n = (size_t)vw->width * vw->height;
    HIP_TRY(hipMemsetAsync(dev_rows, 0, n * (kRL * kRS + 1) * sizeof(float), c->stream));
    int m = 0;
    rc = rpnn_march(c, v, vw, 0, dev_out4, &m);
    if (rc || m == 0) return rc;
    const v3 LX = normalize3(mk3(vw->light_dir[0], vw->light_dir[1], vw->light_dir[2]));
    hipLaunchKernelGGL(k_rpnn_rows, dim3(m), dim3(256), 0, c->stream, mrp_dev_volume(v), c->rpnn->d_recs, m, LX,
                       vw->alpha, dev_rows);
    HIP_TRY(hipGetLastError());
    return MRP_OK;
}
