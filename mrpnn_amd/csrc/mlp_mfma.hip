// MRPNN network on MFMA (gfx950): BF16 (weights as one or two bf16 terms) and FP8 e4m3, FP32
// accumulation.  Layout and
// K-slot order are described in mlp_mfma.h.  The dataflow is RadiancePredictInternal
// (core/radiancePredict.cu:301-422):
//   * 12 SE/SERES blocks (3 branches each) as 16x16x32 MFMAs; gate MLPs (8->8->3), pooling
//     statistics and residual/replace bookkeeping (ADD3/REP3) in FP32 on the VALU;
//   * LSEFin1 (75->16) accumulated incrementally in FP32 as each block's pooled statistics
//     appear, LSEFin2 and the whole tail (128->128->64->32->16, 6 residual pairs, 16->1) on
//     MFMA.
// Feature providers: the fused kernel gathers the 192 stencil points of each pixel straight
// into the B-operand layout (lane (c,q) computes exactly the points it feeds: no LDS
// transposition); the features kernel reads 577-float rows from HBM.
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/mrp_c_api.h"
#include "capi_internal.h"
#include "dev_common.hpp"
#include "mlp_mfma.h"
#include "dev_gather_fast.hpp"
#include "net_layout.h"
#include "fp8_scales.h"

namespace mrp {

// ------------------------------------------------------------------------------------------
// Host: fragment packing
// ------------------------------------------------------------------------------------------
uint16_t host_f32_to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    return (uint16_t)u;
}
uint8_t host_f32_to_e4m3(float f) {  // OCP e4m3fn, round to nearest even, saturating at 448
    uint32_t u;
    memcpy(&u, &f, 4);
    uint8_t sign = (u >> 24) & 0x80;
    double a = fabs((double)f);
    if (a == 0.0) return sign;
    if (a >= 448.0) return sign | 0x7e;
    int e;
    double m = frexp(a, &e);  // a = m * 2^e, m in [0.5,1)
    int E = e - 1;            // a = (2m) * 2^E, 2m in [1,2)
    if (E < -6) {             // subnormal: k * 2^-9
        double k = a * 512.0;
        double r = nearbyint(k);  // ties to even (default rounding mode)
        if (r >= 8.0) return sign | 0x08;
        return sign | (uint8_t)r;
    }
    double mant = (2.0 * m - 1.0) * 8.0;
    double r = nearbyint(mant);
    if (r >= 8.0) {
        r = 0.0;
        ++E;
    }
    if (E > 8 || (E == 8 && r > 6.0)) return sign | 0x7e;
    return sign | (uint8_t)(((E + 7) << 3) | (int)r);
}
// largest e with amax * 2^e <= 448 (exact, from the binary exponent)
__host__ __device__ inline int fp8_exp_for(float amax) {
    if (!(amax > 0.0f)) return 0;
    int x;
    float f = frexpf(amax, &x);  // amax = f * 2^x, f in [0.5,1); 448 = 0.875 * 2^9
    return f <= 0.875f ? 9 - x : 8 - x;
}
int fp8_exp_for_max(float amax) { return fp8_exp_for(amax); }

namespace {
struct Spec {
    int w, b;       // offsets in the fp32 weight blob (b < 0: no bias)
    int M, K;       // rows, inputs of the source matrix
    int roff;       // first destination row inside the padded M-tiles
    bool k32;       // true: K in chunks of 32; false: K <= 16 in one chunk, slots 4..7 zero
};
void layer_specs(Spec* s) {
    using namespace net;
    const int tw[3][9] = {{L01W, L11W, L21W, L31W, L41W, L51W, L61W, L71W, L81W},
                          {L_Tr01W, L_Tr11W, L_Tr21W, L_Tr31W, L_Tr41W, L_Tr51W, L_Tr61W, L_Tr71W, L_Tr81W},
                          {L_Hg01W, L_Hg11W, L_Hg21W, L_Hg31W, L_Hg41W, L_Hg51W, L_Hg61W, L_Hg71W, L_Hg81W}};
    const int tb[3][9] = {{L01B, L11B, L21B, L31B, L41B, L51B, L61B, L71B, L81B},
                          {L_Tr01B, L_Tr11B, L_Tr21B, L_Tr31B, L_Tr41B, L_Tr51B, L_Tr61B, L_Tr71B, L_Tr81B},
                          {L_Hg01B, L_Hg11B, L_Hg21B, L_Hg31B, L_Hg41B, L_Hg51B, L_Hg61B, L_Hg71B, L_Hg81B}};
    const int tn[9] = {8, 8, 16, 16, 16, 32, 32, 32, 32};
    for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 9; ++k) s[ML_TRUNK + 9 * b + k] = {tw[b][k], tb[b][k], tn[k], tn[k], 0, tn[k] == 32};
    const int dw[3][4] = {{LD01W, LD11W, LD21W, LD31W}, {LD_Tr01W, LD_Tr11W, LD_Tr21W, LD_Tr31W},
                          {LD_Hg01W, LD_Hg11W, LD_Hg21W, LD_Hg31W}};
    const int db[3][4] = {{LD01B, LD11B, LD21B, LD31B}, {LD_Tr01B, LD_Tr11B, LD_Tr21B, LD_Tr31B},
                          {LD_Hg01B, LD_Hg11B, LD_Hg21B, LD_Hg31B}};
    for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 4; ++k) s[ML_DI + 4 * b + k] = {dw[b][k], db[b][k], 8, 8, 0, false};
    s[ML_LD41A_VAL] = {LD41W, LD41B, 8, 8, 0, false};
    s[ML_LD41A_TR] = {LD_Tr41W, LD_Tr41B, 8, 8, 8, false};
    s[ML_LD41B_HG] = {LD_Hg41W, LD_Hg41B, 8, 8, 0, false};
    s[ML_FIN2] = {LSEFin2W, -1, 6, 16, 0, false};
    s[ML_LC0] = {LC0W, LC0B, 128, 128, 0, true};
    s[ML_LC1] = {LC1W, LC1B, 64, 128, 0, true};
    s[ML_LC2] = {LC2W, LC2B, 32, 64, 0, true};
    s[ML_LX] = {LXW, LXB, 16, 32, 0, true};
    const int xw[12] = {LX0W, LX1W, LX2W, LX3W, LX4W, LX5W, LX6W, LX7W, LX8W, LX9W, LX10W, LX11W};
    const int xb[12] = {LX0B, LX1B, LX2B, LX3B, LX4B, LX5B, LX6B, LX7B, LX8B, LX9B, LX10B, LX11B};
    for (int k = 0; k < 12; ++k) s[ML_LXK + k] = {xw[k], xb[k], 16, 16, 0, false};
    s[ML_LX12] = {LX12W, LX12B, 1, 16, 0, false};
}
}  // namespace

int mfma_net_create(const float* W, MfmaNet* out, int* wexp) {
    Spec spec[ML_COUNT];
    layer_specs(spec);
    std::vector<uint16_t> bf, bflo;  // 8 per lane per fragment
    std::vector<uint8_t> f8;
    std::vector<float> bias;
    int frag = 0;
    for (int id = 0; id < ML_COUNT; ++id) {
        const Spec& s = spec[id];
        const MfmaShape sh = mfma_shape(id);
        if (sh.M != s.M || sh.K != s.K || sh.roff != s.roff || (sh.k32 != 0) != s.k32 || frag != mfma_frag(id) ||
            (int)bias.size() != mfma_bias(id) || mfma_has_bias(id) != (s.b >= 0))
            return mrp_fail(MRP_ERR_STATE, "MFMA layer table mismatch at layer %d", id);
        const int nm = mfma_nm(id), nk = mfma_nk(id);
        float amax = 0.0f;
        for (int i = 0; i < s.M * s.K; ++i) amax = fmaxf(amax, fabsf(W[s.w + i]));
        wexp[id] = fp8_exp_for(amax);
        const float wsc = ldexpf(1.0f, wexp[id]);
        for (int mt = 0; mt < nm; ++mt)
            for (int kc = 0; kc < nk; ++kc) {
                for (int lane = 0; lane < 64; ++lane) {
                    const int q = lane >> 4, row = 16 * mt + (lane & 15) - s.roff;
                    for (int slot = 0; slot < 8; ++slot) {
                        int k;
                        if (s.k32) k = 32 * kc + (slot < 4 ? 4 * q + slot : 16 + 4 * q + slot - 4);
                        else k = slot < 4 ? 4 * q + slot : -1;
                        float v = (row >= 0 && row < s.M && k >= 0 && k < s.K) ? W[s.w + row * s.K + k] : 0.0f;
                        const uint16_t hi = host_f32_to_bf16(v);
                        const uint32_t hb = (uint32_t)hi << 16;
                        float hf;
                        memcpy(&hf, &hb, 4);
                        bf.push_back(hi);
                        bflo.push_back(host_f32_to_bf16(v - hf));
                        f8.push_back(host_f32_to_e4m3(v * wsc));
                    }
                }
                ++frag;
            }
        if (s.b >= 0)
            for (int r = 0; r < 16 * nm; ++r) {
                int src = r - s.roff;
                bias.push_back(src >= 0 && src < s.M ? W[s.b + src] : 0.0f);
            }
    }
    void *dbf = nullptr, *dbl = nullptr, *df8 = nullptr, *db = nullptr, *ds = nullptr;
    const size_t sbytes = ML_COUNT * sizeof(float4) + ML_COUNT * 4;  // scales, then calibration words
    if (hipMalloc(&dbf, bf.size() * 2) != hipSuccess || hipMalloc(&dbl, bflo.size() * 2) != hipSuccess ||
        hipMalloc(&df8, f8.size()) != hipSuccess || hipMalloc(&db, bias.size() * 4) != hipSuccess ||
        hipMalloc(&ds, sbytes) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "MFMA weight allocation failed");
    if (hipMemcpy(dbf, bf.data(), bf.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dbl, bflo.data(), bflo.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(df8, f8.data(), f8.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, bias.data(), bias.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(ds, 0, sbytes) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "MFMA weight upload failed");
    out->bf16 = (const uint4*)dbf;
    out->bf16_lo = (const uint4*)dbl;
    out->fp8 = (const uint2*)df8;
    out->bias = (const float*)db;
    out->f8 = (const float4*)ds;
    int rc = mfma_net_set_fp8_exps(*out, wexp, kFp8ActExp);
    if (rc) return rc;
    // LDS images of the LDS-weight kernels: hi fragments of every layer, lo fragments of the
    // two-term layers, biases (offsets lds_hi_off / lds_lo_off / lds_frag_bytes)
    const int precs[3] = {MRP_BF16, MRP_FP8, MRP_BF16_1T};
    for (int pi = 0; pi < 3; ++pi) {
        const int v = precs[pi] | kLdsW;
        std::vector<uint8_t> img(lds_weight_bytes(v), 0);
        for (int id = 0; id < ML_COUNT; ++id) {
            const int nf = mfma_nm(id) * mfma_nk(id), fb = frag_bytes(v, id), lb = fb / 64;
            for (int f = 0; f < nf; ++f)
                for (int lane = 0; lane < 64; ++lane) {
                    const size_t g = (size_t)(mfma_frag(id) + f) * 64 + lane;  // 8 values per lane
                    uint8_t* dst = &img[lds_hi_off(v, id) + f * fb + lane * lb];
                    if (precs[pi] == MRP_FP8) {
                        memcpy(dst, &f8[g * 8], 8);
                    } else {
                        memcpy(dst, &bf[g * 8], lb);  // slots 0..3 (compact) or 0..7
                        if (two_term(v, id)) memcpy(&img[lds_lo_off(v, id) + f * fb + lane * lb], &bflo[g * 8], lb);
                    }
                }
        }
        memcpy(&img[lds_frag_bytes(v)], bias.data(), bias.size() * 4);
        void* d = nullptr;
        if (hipMalloc(&d, img.size()) != hipSuccess) return mrp_fail(MRP_ERR_NOMEM, "LDS image allocation failed");
        if (hipMemcpy(d, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess)
            return mrp_fail(MRP_ERR_HIP, "LDS image upload failed");
        out->img[precs[pi]] = (const uint4*)d;
    }
    return MRP_OK;
}

int mfma_net_set_fp8_exps(const MfmaNet& net, const int* wexp, const int* aexp) {
    float4 t[ML_COUNT];
    for (int id = 0; id < ML_COUNT; ++id) {
        if (aexp[id] < -100 || aexp[id] > 100) return mrp_fail(MRP_ERR_INVALID, "FP8 exponent of layer %d out of range", id);
        t[id] = make_float4(ldexpf(1.0f, -aexp[id]), ldexpf(448.0f, -aexp[id]), ldexpf(1.0f, -(aexp[id] + wexp[id])),
                            0.0f);
    }
    if (hipMemcpy((void*)net.f8, t, sizeof t, hipMemcpyHostToDevice) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "FP8 scale upload failed");
    return MRP_OK;
}

void mfma_net_destroy(MfmaNet* net) {
    hipFree((void*)net->bf16);
    hipFree((void*)net->bf16_lo);
    hipFree((void*)net->fp8);
    hipFree((void*)net->bias);
    hipFree((void*)net->f8);
    for (int i = 0; i < 4; ++i) {
        hipFree((void*)net->img[i]);
        net->img[i] = nullptr;
    }
    net->f8 = nullptr;
    net->bf16 = nullptr;
    net->fp8 = nullptr;
    net->bias = nullptr;
}

// ------------------------------------------------------------------------------------------
// Device
// ------------------------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ inline uint32_t pack_bf16x2(float a, float b) {  // v_cvt_pk_bf16_f32, round to nearest even
    const bf16x2_t v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
typedef short s16x2_t __attribute__((ext_vector_type(2)));
// Four values to fp8 e4m3 as x / 2^exponent(scale) (v_cvt_scalef32_pk_fp8_f32: round to nearest
// even, only the scale's exponent is used; a result beyond 448 becomes NaN, so callers clamp).
__device__ inline uint32_t pack_fp8x4_scaled(float a, float b, float c, float d, float scale) {
    s16x2_t w = {0, 0};
    w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(w, a, b, scale, false);
    w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f32(w, c, d, scale, true);
    return __builtin_bit_cast(uint32_t, w);
}
__device__ inline float clamp_f8(float x, float lim) { return __builtin_amdgcn_fmed3f(x, -lim, lim); }

// Cross-row lane exchanges of the tile layout (lane = c + 16q) with the gfx950 permlane
// swaps (plain VALU, no LDS round trip).  permlane16_swap(x, x) returns rows (0,0,2,2) and
// (1,1,3,3) of x; permlane32_swap(x, x) returns halves (lo,lo) and (hi,hi).
__device__ inline float q_sum(float x) {  // sum over the 4 lanes of column c, same order as xor 16 then 32
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const float s = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}
__device__ inline float q_max(float x) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const float s = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
    return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}
__device__ inline float xor32(float x, int lane) {  // value of lane ^ 32
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(lane < 32 ? b[1] : b[0]);
}

// Weight and bias fragments through buffer resources: the per-layer offset is a compile-time
// scalar offset and the lane offset a fixed VGPR, so a fragment load needs no address
// arithmetic.
struct NetRs {
    BufRsrc w;    // bf16 hi (16 bytes per lane per fragment) or fp8 (8 bytes) fragments
    BufRsrc wlo;  // bf16 lo fragments
    BufRsrc bias;
    const float4* f8;  // FP8 per-layer scales (MfmaNet::f8)
    unsigned* cal;     // kCal: per-layer max |input|
    bool cal_lane;     // kCal: the lane's pixel is a real row (padding lanes are not recorded)
    const char* lw;    // LDS copy of the weight image (variants with kLdsW)
};
// The pointers are uniform (kernel arguments); reading them through readfirstlane keeps the
// descriptors in scalar registers also where they arrive in vector registers (call arguments).
__device__ inline void* uniform_ptr(const void* p) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (void*)(((uint64_t)hi << 32) | lo);
}
__device__ inline NetRs make_net_rs(const MfmaNet& n, bool fp8, bool lane_ok) {
    NetRs r;
    r.w = fp8 ? __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(n.fp8), (short)0, kMfmaFrags * 512, 0x00020000)
              : __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(n.bf16), (short)0, kMfmaFrags * 1024, 0x00020000);
    r.wlo = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(n.bf16_lo), (short)0, kMfmaFrags * 1024, 0x00020000);
    r.bias = __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(n.bias), (short)0, kMfmaBias * 4, 0x00020000);
    r.f8 = (const float4*)uniform_ptr(n.f8);
    r.cal = (unsigned*)uniform_ptr(mfma_cal_words(n));
    r.cal_lane = lane_ok;
    r.lw = (const char*)n.lds;
    return r;
}

// A bf16 weight fragment of one lane from the LDS image: 16 bytes, or 8 (slots 0..3) for a
// compact fragment whose slots 4..7 are zero.
template <bool CMP>
__device__ inline u32x4_t lds_frag(const char* p) {
    if constexpr (CMP) {
        const uint2 v = *(const uint2*)p;
        return u32x4_t{v.x, v.y, 0u, 0u};
    } else {
        return *(const u32x4_t*)p;
    }
}

// One bf16 MFMA step acc += W x: a K <= 16 layer (CMP: compact fragments, inputs in slots 0..3 of
// each lane) runs as a 16x16x16 MFMA on those slots (the same products: the 16 x 32 form's slots
// 4..7 are zero in both operands), a wider one as 16x16x32.
typedef short s16x4_t __attribute__((ext_vector_type(4)));
template <bool CMP>
__device__ inline f32x4_t mfma_bf16(const u32x4_t& w, const u32x4_t& x, const f32x4_t& acc) {
    if constexpr (CMP) {
        typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
        const s16x4_t w4 = __builtin_bit_cast(s16x4_t, u32x2_t{w[0], w[1]});
        const s16x4_t x4 = __builtin_bit_cast(s16x4_t, u32x2_t{x[0], x[1]});
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w4, x4, acc, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, w), __builtin_bit_cast(bf16x8_t, x),
                                                       acc, 0, 0, 0);
    }
}

// Calibration: the largest |x| over one layer call's inputs (the real rows of the 16-pixel tile),
// atomically folded into net.cal[id] (non-negative floats order as their bit patterns).
template <int N>
__device__ inline void cal_note(const NetRs& net, int id, const float* x) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) m = fmaxf(m, fabsf(x[i]));
    if (!net.cal_lane) m = 0.0f;
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(net.cal + id, __float_as_uint(m));
}

// out[NM][4] = act(W_ID * in + b): in[NK][8] in the K-slot order of mlp_mfma.h.  One M-tile
// at a time behind a scheduling barrier, so at most NK weight fragments are in flight.
template <int PREC, int ID, bool SB = true>
__device__ inline void dense(const NetRs& net, const float (&in)[mfma_nk(ID)][8], float (&out)[mfma_nm(ID)][4],
                             int lane, bool relu) {
    constexpr int NM = mfma_nm(ID), NK = mfma_nk(ID), FRAG = mfma_frag(ID), BIAS = mfma_bias(ID);
    // second weight term: MRP_BF16 applies it to the tail (LSEFin2, LC0..LX12), where the
    // single-term rounding error concentrates; the trunk's 8..32-wide layers use one term
    constexpr int P = base_prec(PREC);
    constexpr bool TWO = two_term(PREC, ID);
    // LDS image offsets of this layer (constexpr variables: the table walks fold at compile
    // time); compact fragments hold slots 0..3 of each lane (8 bytes)
    constexpr bool LDS = lds_weights(PREC);
    constexpr int HO = lds_hi_off(PREC, ID), LO = lds_lo_off(PREC, ID), FB = frag_bytes(PREC, ID);
    constexpr bool CMP = compact_frag(PREC, ID);
    constexpr int BO = lds_frag_bytes(PREC) + BIAS * 4;
    const int q = lane >> 4;
    // Fragments of M-tile mt+1 are requested before the MFMAs of M-tile mt (two buffers), so
    // the weight latency of a multi-tile layer is paid once.
    if constexpr (P == MRP_BF16 || P == MRP_BF16_1T) {
        if constexpr (calibrating(PREC)) cal_note<NK * 8>(net, ID, &in[0][0]);
        u32x4_t b[NK];
#pragma unroll
        for (int kc = 0; kc < NK; ++kc)
            b[kc] = u32x4_t{pack_bf16x2(in[kc][0], in[kc][1]), pack_bf16x2(in[kc][2], in[kc][3]),
                            pack_bf16x2(in[kc][4], in[kc][5]), pack_bf16x2(in[kc][6], in[kc][7])};
        u32x4_t wa[2][NK], wl[2][NK], bb[2];
        auto fetch = [&](int mt, int buf) {
#pragma unroll
            for (int kc = 0; kc < NK; ++kc) {
                if constexpr (LDS) {
                    const int fo = (mt * NK + kc) * FB + lane * (FB / 64);
                    wa[buf][kc] = lds_frag<CMP>(net.lw + HO + fo);
                    if constexpr (TWO) wl[buf][kc] = lds_frag<CMP>(net.lw + LO + fo);
                } else {
                    const int fo = (FRAG + mt * NK + kc) * 1024;
                    wa[buf][kc] = __builtin_amdgcn_raw_buffer_load_b128(net.w, lane * 16, fo, 0);
                    if constexpr (TWO) wl[buf][kc] = __builtin_amdgcn_raw_buffer_load_b128(net.wlo, lane * 16, fo, 0);
                }
            }
            bb[buf] = u32x4_t{0u, 0u, 0u, 0u};
            if constexpr (mfma_has_bias(ID)) {
                if constexpr (LDS)
                    bb[buf] = *(const u32x4_t*)(net.lw + BO + 16 * mt * 4 + q * 16);
                else
                    bb[buf] = __builtin_amdgcn_raw_buffer_load_b128(net.bias, q * 16, (BIAS + 16 * mt) * 4, 0);
            }
        };
        fetch(0, 0);
        // the first M-tile's loads (both weight terms and the bias) are issued together:
        // otherwise register reuse can serialise the second term's load behind the first MFMA
        if constexpr (SB) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < NM; ++mt) {
            const int cur = mt & 1;
            if (mt + 1 < NM) fetch(mt + 1, cur ^ 1);
            // bias first, then the products (the reference's LinearReLU_SFL order); the lo terms
            // accumulate in a second, independent chain added at the end, so the two MFMA chains
            // overlap instead of forming one chain twice as long
            f32x4_t acc = f32x4_t{__uint_as_float(bb[cur][0]), __uint_as_float(bb[cur][1]),
                                  __uint_as_float(bb[cur][2]), __uint_as_float(bb[cur][3])};
            f32x4_t acl = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < NK; ++kc) {
                acc = mfma_bf16<CMP>(wa[cur][kc], b[kc], acc);
                if constexpr (TWO) acl = mfma_bf16<CMP>(wl[cur][kc], b[kc], acl);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float v = TWO ? acc[i] + acl[i] : acc[i];
                out[mt][i] = relu ? fmaxf(v, 0.0f) : v;
            }
            if constexpr (SB) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        // per-layer max-abs scales: inputs clamped to +-448 * 2^-aexp, stored as x * 2^aexp
        const float4 s8 = net.f8[ID];
        long b[NK];
#pragma unroll
        for (int kc = 0; kc < NK; ++kc) {
            float v[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) v[s] = clamp_f8(in[kc][s], s8.y);
            const uint32_t lo = pack_fp8x4_scaled(v[0], v[1], v[2], v[3], s8.x);
            const uint32_t hi = pack_fp8x4_scaled(v[4], v[5], v[6], v[7], s8.x);
            b[kc] = (long)(((uint64_t)hi << 32) | lo);
        }
        const float ds = s8.z;
        long wa[2][NK];
        u32x4_t bb[2];
        auto fetch = [&](int mt, int buf) {
#pragma unroll
            for (int kc = 0; kc < NK; ++kc) {
                if constexpr (LDS)
                    wa[buf][kc] = *(const long*)(net.lw + HO + (mt * NK + kc) * 512 + lane * 8);
                else
                    wa[buf][kc] = __builtin_bit_cast(
                        long, __builtin_amdgcn_raw_buffer_load_b64(net.w, lane * 8, (FRAG + mt * NK + kc) * 512, 0));
            }
            bb[buf] = u32x4_t{0u, 0u, 0u, 0u};
            if constexpr (mfma_has_bias(ID)) {
                if constexpr (LDS)
                    bb[buf] = *(const u32x4_t*)(net.lw + BO + 16 * mt * 4 + q * 16);
                else
                    bb[buf] = __builtin_amdgcn_raw_buffer_load_b128(net.bias, q * 16, (BIAS + 16 * mt) * 4, 0);
            }
        };
        fetch(0, 0);
        if constexpr (SB) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < NM; ++mt) {
            const int cur = mt & 1;
            if (mt + 1 < NM) fetch(mt + 1, cur ^ 1);
            f32x4_t acc = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < NK; ++kc) acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(wa[cur][kc], b[kc], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = fmaf(acc[i], ds, __uint_as_float(bb[cur][i]));
                out[mt][i] = relu ? fmaxf(v, 0.0f) : v;
            }
            if constexpr (SB) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Makes a pointer opaque to the optimiser once per tile iteration: the network's weights are
// loop-invariant in the persistent tile loop, and without this the compiler hoists all of
// their loads (hundreds of registers) above the loop and spills.  The pointer is offset by a
// zero the compiler cannot see through, so it keeps its address space (global loads for
// kernel arguments, LDS loads for shared arrays).
__device__ inline int opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}
template <class T>
__device__ inline T* opaque(T* p) {
    return p + opaque_zero();
}
template <class T>
__device__ inline T* opaque_v(T* p) {
    return p + opaque_zero();
}

// Gate sigmoid with the hardware exp2 and reciprocal (v_rcp_f32, 1 ulp; about 2 ulp in all):
// the gates feed BF16 / FP8 products whose rounding is far coarser.
__device__ inline float sigmoidf_(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// Per-pixel lane state shared by the tile routine.
struct LaneCtx {
    int lane, c, q;
    bool ok;        // this column holds a pixel
    float g, gamma;
    v3 sr;          // scatterrate passed to the network (scatter_rate / 1.001)
    v3 sp;          // sr^4 (radiancePredict.cu:324)
    bool need_y;    // sr.y != sr.x: channel y needs its own tail
    bool need_z;    // sr.z differs from sr.x and sr.y
};

// SE gate weights of the 12 gated blocks (trunk blocks 0..7, DI blocks 0..3) staged in LDS:
// per gate 64 floats SEW0 (8x8, row-major) then 24 floats SEW1 (3x8).
constexpr int kGates = 12, kGateFloats = 88;
__device__ inline int gate_w0(int gi) {
    using namespace net;
    constexpr int w0[kGates] = {LSE01W, LSE11W, LSE21W, LSE31W, LSE41W, LSE51W, LSE61W, LSE71W,
                                LDSE01W, LDSE11W, LDSE21W, LDSE31W};
    return w0[gi];
}
__device__ inline int gate_w1(int gi) {
    using namespace net;
    constexpr int w1[kGates] = {LSE02W, LSE12W, LSE22W, LSE32W, LSE42W, LSE52W, LSE62W, LSE72W,
                                LDSE02W, LDSE12W, LDSE22W, LDSE32W};
    return w1[gi];
}

// x0..x3 = value of x in rows 0..3 of the lane's column (lane = c + 16 row), in every lane.
__device__ inline void q_gather(float x, float& x0, float& x1, float& x2, float& x3) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const auto a = __builtin_amdgcn_permlane32_swap(p[0], p[0], false, false);  // rows (0,0,2,2)
    const auto b = __builtin_amdgcn_permlane32_swap(p[1], p[1], false, false);  // rows (1,1,3,3)
    x0 = __uint_as_float(a[0]);
    x2 = __uint_as_float(a[1]);
    x1 = __uint_as_float(b[0]);
    x3 = __uint_as_float(b[1]);
}

// Pooled statistics and SE gate of one block (radiancePredict.cu:240-266): mean/max of the
// raw slice of each branch, gate MLP SEW0 (8x8, ReLU) and SEW1 (3x8, sigmoid); the six
// pooled values are folded into the LSEFin1 accumulators (rows 4q+i of this lane).  The
// four lanes of a pixel split the gate: lane q computes hidden units 2q, 2q+1 and gate
// output min(q, 2); the values are exchanged with permlane swaps.  Every unit is the same
// FMA chain as the scalar form, so the gate values are unchanged.
__device__ inline void block_gate(const float* __restrict__ sG, const float* __restrict__ sW1, const LaneCtx& lc,
                                  int pid, int gi, const float (&F)[3][8], int nv, bool lane_has, int n, float (&s)[3],
                                  float (&hacc)[4]) {
    float sum[3], mx[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        float a = 0.0f, m = -10000.0f;
        if (lane_has) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < nv) {
                    a += F[b][i];
                    m = fmaxf(m, F[b][i]);
                }
        }
        a = q_sum(a);
        m = q_max(m);
        sum[b] = a * (1.0f / (float)n);  // n is a power of two: exact
        mx[b] = m;
    }
    const float t[8] = {sum[0], sum[1], sum[2], mx[0], mx[1], mx[2], lc.g, lc.gamma};
    const float* g = sG + gi * kGateFloats;
    const float* w0 = g + 16 * lc.q;  // rows 2q, 2q+1 of SEW0
    float ha = 0.0f, hb = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        ha = fmaf(t[j], w0[j], ha);
        hb = fmaf(t[j], w0[8 + j], hb);
    }
    float h[8];
    q_gather(fmaxf(ha, 0.0f), h[0], h[2], h[4], h[6]);
    q_gather(fmaxf(hb, 0.0f), h[1], h[3], h[5], h[7]);
    const float* w1 = g + 64 + 8 * min(lc.q, 2);
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) acc = fmaf(h[j], w1[j], acc);
    float s3;
    q_gather(sigmoidf_(acc), s[0], s[1], s[2], s3);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* wr = sW1 + (4 * lc.q + i) * 75;
        float a = hacc[i];
#pragma unroll
        for (int b = 0; b < 3; ++b) a = fmaf(wr[pid + b], sum[b], a);
#pragma unroll
        for (int b = 0; b < 3; ++b) a = fmaf(wr[36 + pid + b], mx[b], a);
        hacc[i] = a;
    }
}

// Feature providers fill F[b][slot] for points p0+4q+i (slots 0..3) and, when wide,
// p0+16+4q+i (slots 4..7); b = 0 X_Val, 1 X_Val_Sub, 2 X_Val_Hg.
struct GatherSrc {
    const DevVolume* vol;
    const StencilEntry* st;  // LDS copy
    PixelState ps;
    float g, alpha;
    __device__ void get(const LaneCtx& lc, int p0, bool wide, float (&F)[3][8]) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !wide) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = p0 + 16 * h + 4 * lc.q + i;
                v3 f = lc.ok ? gather_point(*vol, st[k], k, ps, g, alpha) : mk3(0.0f, 0.0f, 0.0f);
                F[0][4 * h + i] = f.x;
                F[2][4 * h + i] = f.y;
                F[1][4 * h + i] = f.z;
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (!wide)
#pragma unroll
            for (int i = 4; i < 8; ++i) F[0][i] = F[1][i] = F[2][i] = 0.0f;
    }
};
// Fast gather (dev_gather_fast.hpp) from the zero-padded pyramid copies.
struct FastGatherSrc {
    const FastStencil* st;  // LDS copy
    FastView fv;
    FastPixel px;
    __device__ void get(const LaneCtx& lc, int p0, bool wide, float (&F)[3][8]) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !wide) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = p0 + 16 * h + 4 * lc.q + i;
                v3 f = mk3(0.0f, 0.0f, 0.0f);
                // chunks 16..159: one density mip per chunk, known here (gather_chunk); chunk 0
                // and the light points mix two mips across the lane rows (table geometry).
                // Density and TR levels share the resolution for stencil mips >= 1: all points
                // of chunks 16..159 and the second half of the light points
                const bool same = (p0 >= 16 && p0 < 160) || (p0 == 160 && h == 1);
                // the lane test also keeps the compiler from interleaving consecutive points
                // (without it the kernel spills)
                if (!lc.ok) {
                } else if (p0 == 16) f = gather_chunk<2>(fv, st[k], px);
                else if (p0 == 32) f = gather_chunk<3>(fv, st[k], px);
                else if (p0 == 48) f = gather_chunk<4>(fv, st[k], px);
                else if (p0 == 64) f = gather_chunk<5>(fv, st[k], px);
                else if (p0 == 96) f = gather_chunk<6>(fv, st[k], px);
                else if (p0 == 128) f = gather_chunk<7>(fv, st[k], px);
                else if (p0 >= 160) f = gather_fast<true>(fv, st[k], px, false, same);
                else f = gather_fast<false>(fv, st[k], px, p0 == 0 && h == 0 && i == 0 && lc.q == 0, same);
                F[0][4 * h + i] = f.x;
                F[2][4 * h + i] = f.y;
                F[1][4 * h + i] = f.z;
                // two points' loads in flight at a time: more hides memory latency, fewer
                // registers keep the kernel at two waves per SIMD
                if (i == 3) __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (!wide)
#pragma unroll
            for (int i = 4; i < 8; ++i) F[0][i] = F[1][i] = F[2][i] = 0.0f;
    }
};
// The fast gather of one pixel (record state ps) of a view (constants vc, anisotropy g).
__device__ inline FastGatherSrc fast_src(const DevVolume& vol, const ViewConst& vc, float g, const PixelState& ps,
                                         const FastStencil* sfs, const float* lut_rows) {
    FastGatherSrc src;
    src.st = sfs;
    src.fv.qm = make_rsrc((const float*)vol.qmips, 4 * kQuadTexels);
    src.fv.qt = make_rsrc(vol.ptr, 4 * kQuadTrTexels);
    src.fv.rows = make_rsrc(lut_rows, (long long)kPC * kLutRow);
    src.fv.L = ps.L;
    src.fv.LX = ps.LX;
    src.fv.invLX = vc.invLX;
    src.fv.a2f = mk3(0.4999f, 0.4999f, 0.4999f) * vc.invLX;
    src.fv.alpha64 = vc.alpha64;
    src.px.pos = ps.pos;
    src.px.X = ps.X;
    src.px.Y = ps.Y;
    src.px.Z = ps.Z;
    src.px.ax = dot3(ps.LX, ps.X);
    src.px.ay = dot3(ps.LX, ps.Y);
    src.px.az = dot3(ps.LX, ps.Z);
    src.px.uc1 = fmaf(dot3(ps.X, ps.LX), 0.5f, 0.5f);
    src.px.uc2 = fmaf(dot3(ps.LX, ps.LX), 0.5f, 0.5f);
    src.px.xh_centre = log_ge1(hg_fast(dot3(ps.X, ps.L), g) + 1.0f);
    return src;
}

struct RowSrc {
    const float* row;  // 577 floats [X_Val | X_Val_Hg | X_Val_Sub | gamma]
    __device__ void get(const LaneCtx& lc, int p0, bool wide, float (&F)[3][8]) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !wide) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = p0 + 16 * h + 4 * lc.q + i;
                F[0][4 * h + i] = lc.ok ? row[k] : 0.0f;
                F[2][4 * h + i] = lc.ok ? row[kPC + k] : 0.0f;
                F[1][4 * h + i] = lc.ok ? row[2 * kPC + k] : 0.0f;
            }
        }
        if (!wide)
#pragma unroll
            for (int i = 4; i < 8; ++i) F[0][i] = F[1][i] = F[2][i] = 0.0f;
    }
};

// One SE/SERES block for all three branches: in = F*s (+ R), out = ReLU(W in + b).
// Layer ids: ID0 + STRIDE*branch.
template <int PREC, int ID0, int STRIDE>
__device__ inline void block_dense(const NetRs& net, const LaneCtx& lc, const float (&F)[3][8], const float (&s)[3],
                                   const float (&R)[3][8], bool has_res, bool lane_has, float (&out)[3][8]) {
    constexpr int NM = mfma_nm(ID0);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        float in[1][8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            // one rounding (the sum is rounded to bf16 / fp8 for the MFMA right after)
            const float v = has_res ? fmaf(F[b][i], s[b], R[b][i]) : F[b][i] * s[b];
            in[0][i] = (lane_has && (NM == 2 || i < 4)) ? v : 0.0f;
        }
        float o[NM][4];
        // the three branches are independent: no scheduling barrier between them, so their
        // weight loads and MFMAs overlap
        if (b == 0) dense<PREC, ID0, false>(net, in, o, lc.lane, true);
        else if (b == 1) dense<PREC, ID0 + STRIDE, false>(net, in, o, lc.lane, true);
        else dense<PREC, ID0 + 2 * STRIDE, false>(net, in, o, lc.lane, true);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            out[b][i] = o[0][i];
            out[b][4 + i] = NM == 2 ? o[NM - 1][i] : 0.0f;
        }
    }
    __builtin_amdgcn_sched_barrier(0);
}

// Gate statistics + SE/SERES of a 16- or 32-wide trunk block with ADD3 (or REP3 for block 5).
template <int PREC, int K>  // K = 2..7: blocks 3..8 of the main stack
__device__ inline void trunk_block(const NetRs& net, const float* __restrict__ sG, const float* __restrict__ sW1,
                                   const LaneCtx& lc, const float (&F)[3][8], float (&R)[3][8], float (&hacc)[4]) {
    using namespace net;
    constexpr bool wide = K >= 5;
    float s[3], O[3][8];
    block_gate(sG, sW1, lc, 3 * K, K, F, wide ? 8 : 4, true, wide ? 32 : 16, s, hacc);
    block_dense<PREC, ML_TRUNK + K, 9>(net, lc, F, s, R, true, true, O);
    if (K == 4) {  // REP3(32,48,16): A2[32:64] = [out5, X[48:64]]
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                R[b][i] = O[b][i];
                R[b][4 + i] = F[b][i];
            }
    } else {  // ADD3
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int i = 0; i < 8; ++i) R[b][i] = (wide || i < 4) ? O[b][i] + F[b][i] : 0.0f;
    }
}

// One block of the DI stack (K = 0..3), features G already moved to lanes q < 2.
template <int PREC, int K>
__device__ inline void di_block(const NetRs& net, const float* __restrict__ sG, const float* __restrict__ sW1,
                                const LaneCtx& lc, const float (&G)[3][8], float (&R)[3][8], float (&hacc)[4]) {
    using namespace net;
    const bool lo = lc.q < 2;
    float s[3], O[3][8];
    block_gate(sG, sW1, lc, 24 + 3 * K, 8 + K, G, 4, lo, 8, s, hacc);
    block_dense<PREC, ML_DI + K, 4>(net, lc, G, s, R, K > 0, lo, O);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) R[b][i] = (lo && i < 4) ? O[b][i] + G[b][i] : 0.0f;
}

// Fragments of a layer with one M-tile and one K-chunk of at most 16 inputs (the LX chain),
// loaded ahead of the layer that uses them.  FP8 keeps its 8 weight bytes in w[0..1].
struct Frag1 {
    u32x4_t w, wl, b;
};
template <int PREC, int ID>
__device__ inline void frag1_fetch(const NetRs& net, int lane, Frag1& f) {
    static_assert(mfma_nm(ID) == 1 && mfma_nk(ID) == 1 && !mfma_shape(ID).k32, "single-fragment layer");
    constexpr int P = base_prec(PREC);
    constexpr bool TWO = two_term(PREC, ID);
    constexpr int FRAG = mfma_frag(ID), BIAS = mfma_bias(ID);
    constexpr bool LDS = lds_weights(PREC);
    constexpr int HO = lds_hi_off(PREC, ID), LO = lds_lo_off(PREC, ID), FB = frag_bytes(PREC, ID);
    constexpr bool CMP = compact_frag(PREC, ID);
    constexpr int BO = lds_frag_bytes(PREC) + BIAS * 4;
    if constexpr (P == MRP_FP8) {
        if constexpr (LDS) {
            const uint2 v = *(const uint2*)(net.lw + HO + lane * 8);
            f.w = u32x4_t{v.x, v.y, 0u, 0u};
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b64(net.w, lane * 8, FRAG * 512, 0);
            f.w = u32x4_t{v[0], v[1], 0u, 0u};
        }
    } else if constexpr (LDS) {
        f.w = lds_frag<CMP>(net.lw + HO + lane * (FB / 64));
        if constexpr (TWO) f.wl = lds_frag<CMP>(net.lw + LO + lane * (FB / 64));
    } else {
        f.w = __builtin_amdgcn_raw_buffer_load_b128(net.w, lane * 16, FRAG * 1024, 0);
        if constexpr (TWO) f.wl = __builtin_amdgcn_raw_buffer_load_b128(net.wlo, lane * 16, FRAG * 1024, 0);
    }
    if constexpr (LDS)
        f.b = *(const u32x4_t*)(net.lw + BO + (lane >> 4) * 16);
    else
        f.b = __builtin_amdgcn_raw_buffer_load_b128(net.bias, (lane >> 4) * 16, BIAS * 4, 0);
}
// out = act(W in + b) for such a layer; in = the 4 rows the lane holds (K slots 0..3), the
// same arithmetic as dense().
template <int PREC, int ID>
__device__ inline void frag1_apply(const NetRs& net, const Frag1& f, const float (&in)[4], float (&out)[4],
                                   bool relu) {
    constexpr int P = base_prec(PREC);
    constexpr bool TWO = two_term(PREC, ID);
    if constexpr (P == MRP_BF16 || P == MRP_BF16_1T) {
        if constexpr (calibrating(PREC)) cal_note<4>(net, ID, in);
        const u32x4_t b = u32x4_t{pack_bf16x2(in[0], in[1]), pack_bf16x2(in[2], in[3]), 0u, 0u};
        f32x4_t acc = f32x4_t{__uint_as_float(f.b[0]), __uint_as_float(f.b[1]), __uint_as_float(f.b[2]),
                              __uint_as_float(f.b[3])};
        constexpr bool CMP = compact_frag(PREC, ID);
        acc = mfma_bf16<CMP>(f.w, b, acc);
        f32x4_t acl = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};  // lo terms: an independent MFMA (as dense())
        if constexpr (TWO) acl = mfma_bf16<CMP>(f.wl, b, acl);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = TWO ? acc[i] + acl[i] : acc[i];
            out[i] = relu ? fmaxf(v, 0.0f) : v;
        }
    } else {
        const float4 s8 = net.f8[ID];
        const long b = (long)(uint64_t)pack_fp8x4_scaled(clamp_f8(in[0], s8.y), clamp_f8(in[1], s8.y),
                                                         clamp_f8(in[2], s8.y), clamp_f8(in[3], s8.y), s8.x);
        const long w = (long)(((uint64_t)f.w[1] << 32) | f.w[0]);
        const float ds = s8.z;
        const f32x4_t acc =
            __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(w, b, f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = fmaf(acc[i], ds, __uint_as_float(f.b[i]));
            out[i] = relu ? fmaxf(v, 0.0f) : v;
        }
    }
}
// Layer ID on fragments f, then f <- fragments of layer NEXT (two layers ahead).
template <int PREC, int ID, int NEXT>
__device__ inline void frag1_step(const NetRs& net, int lane, Frag1& f, const float (&in)[4], float (&out)[4]) {
    frag1_apply<PREC, ID>(net, f, in, out, true);
    if constexpr (NEXT >= 0) frag1_fetch<PREC, NEXT>(net, lane, f);
    __builtin_amdgcn_sched_barrier(0);
}
// The residual chain after LX (radiancePredict.cu:392-409): six pairs
// x += ReLU(LX_{2k+1} ReLU(LX_{2k} x)), then LX12.  Returns LX12's output row 0.
template <int PREC>
__device__ inline float lx_chain(const NetRs& net, int lane, float (&x)[4]) {
    Frag1 fa, fb;
    frag1_fetch<PREC, ML_LXK + 0>(net, lane, fa);
    frag1_fetch<PREC, ML_LXK + 1>(net, lane, fb);
    __builtin_amdgcn_sched_barrier(0);
    float y[4], z[4];
#define MRP_LX_PAIR(K, NA, NB)                                         \
    frag1_step<PREC, ML_LXK + 2 * (K), NA>(net, lane, fa, x, y);        \
    frag1_step<PREC, ML_LXK + 2 * (K) + 1, NB>(net, lane, fb, y, z);    \
    for (int i = 0; i < 4; ++i) x[i] = x[i] + z[i];
    MRP_LX_PAIR(0, ML_LXK + 2, ML_LXK + 3)
    MRP_LX_PAIR(1, ML_LXK + 4, ML_LXK + 5)
    MRP_LX_PAIR(2, ML_LXK + 6, ML_LXK + 7)
    MRP_LX_PAIR(3, ML_LXK + 8, ML_LXK + 9)
    MRP_LX_PAIR(4, ML_LXK + 10, ML_LXK + 11)
    MRP_LX_PAIR(5, ML_LX12, -1)
#undef MRP_LX_PAIR
    float p[4];
    frag1_step<PREC, ML_LX12, -1>(net, lane, fa, x, p);
    return p[0];  // row 0: lanes q == 0
}

// Tail SC(p) for one channel (radiancePredict.cu:375-409); returns P in lanes q == 0.
template <int PREC>
__device__ __attribute__((always_inline)) inline float tail(const NetRs& net, const float* __restrict__ W,
                                                             const float* __restrict__ sW1, const LaneCtx& lc,
                                                             float srp, const float (&hacc)[4],
                                                             const float (&comb)[3][8], const float (&tA)[4],
                                                             const float (&tB)[4]) {
    using namespace net;
    const int q = lc.q;
    const bool lo = q < 2;
    // LSEFin1 rows 4q+i: add the g, gamma, srp columns, ReLU
    float hh[1][8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* wr = sW1 + (4 * q + i) * 75;
        float a = fmaf(wr[72], lc.g, hacc[i]);
        a = fmaf(wr[73], lc.gamma, a);
        a = fmaf(wr[74], srp, a);
        hh[0][i] = fmaxf(a, 0.0f);
        hh[0][4 + i] = 0.0f;
    }
    float d6[1][4];
    dense<PREC, ML_FIN2>(net, hh, d6, lc.lane, false);
    float w6[6];
    {
        float sgm[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) sgm[i] = sigmoidf_(d6[0][i]);
        const int c0 = lc.c, c1 = lc.c + 16;
        w6[0] = __shfl(sgm[0], c0);
        w6[1] = __shfl(sgm[1], c0);
        w6[2] = __shfl(sgm[2], c0);
        w6[3] = __shfl(sgm[3], c0);
        w6[4] = __shfl(sgm[0], c1);
        w6[5] = __shfl(sgm[1], c1);
    }
    // LGGS rows 0..7 into tile B rows 8..15 (lanes q >= 2), bias last (:377)
    float tBc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!lo) {
            const int r = 4 * (q - 2) + i;
            float a = 0.0f;
            a = fmaf(lc.g, W[LGGSW + 3 * r + 0], a);
            a = fmaf(lc.gamma, W[LGGSW + 3 * r + 1], a);
            a = fmaf(srp, W[LGGSW + 3 * r + 2], a);
            a += W[LGGSB + r];
            tBc[i] = fmaxf(a, 0.0f);
        } else {
            tBc[i] = tB[i];
        }
    }
    float in[4][8];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) in[b][i] = comb[b][i] * w6[b];
    const float sa = lo ? w6[3] : w6[4], sb = lo ? w6[5] : 1.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        in[3][i] = tA[i] * sa;
        in[3][4 + i] = tBc[i] * sb;
    }
    float o0[8][4];
    dense<PREC, ML_LC0>(net, in, o0, lc.lane, true);
    float in1[4][8];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            in1[k][i] = o0[2 * k][i];
            in1[k][4 + i] = o0[2 * k + 1][i];
        }
    float o1[4][4];
    dense<PREC, ML_LC1>(net, in1, o1, lc.lane, true);
    float in2[2][8];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            in2[k][i] = o1[2 * k][i];
            in2[k][4 + i] = o1[2 * k + 1][i];
        }
    float o2[2][4];
    dense<PREC, ML_LC2>(net, in2, o2, lc.lane, true);
    float in3[1][8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        in3[0][i] = o2[0][i];
        in3[0][4 + i] = o2[1][i];
    }
    float x[1][4];
    dense<PREC, ML_LX>(net, in3, x, lc.lane, true);
    return lx_chain<PREC>(net, lc.lane, x[0]);
}

// Out-of-line tail for the second and third distinct channels.  The call also shapes the
// kernel's register allocation: with it, the inputs it takes by reference stay in scratch
// and the trunk keeps to the three-waves-per-SIMD register budget without spills.
template <int PREC>
__device__ __attribute__((noinline)) float tail_call(MfmaNet nt, const float* __restrict__ W,
                                                     const float* __restrict__ sW1, const LaneCtx& lc, float srp,
                                                     const float (&hacc)[4], const float (&comb)[3][8],
                                                     const float (&tA)[4], const float (&tB)[4]) {
    return tail<PREC>(make_net_rs(nt, base_prec(PREC) == MRP_FP8, lc.ok), W, sW1, lc, srp, hacc, comb, tA, tB);
}

// Full network on one 16-pixel tile; returns, in lanes q == 0, max(e^P - 1, 0) * srp per
// channel for pixel c.
template <int PREC, class Src>
__device__ inline v3 run_tile(const MfmaNet& nt, const float* __restrict__ W, const float* __restrict__ sG,
                              const float* __restrict__ sW1,
                              const LaneCtx& lc, const Src& src) {
    using namespace net;
    const NetRs net = make_net_rs(nt, base_prec(PREC) == MRP_FP8, lc.ok);
    const bool lo = lc.q < 2;
    float F[3][8], G[3][8], R[3][8], O[3][8], s[3];
    float hacc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // ---- block 1: SE(0,0,8), ADD3(0,0,8); points 0..15 gathered, [0,8) in lanes q<2
    src.get(lc, 0, false, F);
    block_gate(sG, sW1, lc, 0, 0, F, 4, lo, 8, s, hacc);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) R[b][i] = 0.0f;
    block_dense<PREC, ML_TRUNK + 0, 9>(net, lc, F, s, R, false, lo, O);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) R[b][i] = (lo && i < 4) ? O[b][i] + F[b][i] : 0.0f;
    // ---- block 2: SERES(0,8,8,8), REP3(0,8,8); [8,16) lives in lanes q>=2: move down
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) G[b][i] = i < 4 ? xor32(F[b][i], lc.lane) : 0.0f;
    block_gate(sG, sW1, lc, 3, 1, G, 4, lo, 8, s, hacc);
    block_dense<PREC, ML_TRUNK + 1, 9>(net, lc, G, s, R, true, lo, O);
    // A2[0:16] = [out2, X[8:16]]: lanes q<2 hold out2 rows 4q+i, lanes q>=2 their own X[8+4(q-2)+i]
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) R[b][i] = i < 4 ? (lo ? O[b][i] : F[b][i]) : 0.0f;
    // ---- blocks 3..8
    src.get(lc, 16, false, F);
    trunk_block<PREC, 2>(net, sG, sW1, lc, F, R, hacc);
    src.get(lc, 32, false, F);
    trunk_block<PREC, 3>(net, sG, sW1, lc, F, R, hacc);
    src.get(lc, 48, false, F);
    trunk_block<PREC, 4>(net, sG, sW1, lc, F, R, hacc);
    src.get(lc, 64, true, F);
    trunk_block<PREC, 5>(net, sG, sW1, lc, F, R, hacc);
    src.get(lc, 96, true, F);
    trunk_block<PREC, 6>(net, sG, sW1, lc, F, R, hacc);
    src.get(lc, 128, true, F);
    trunk_block<PREC, 7>(net, sG, sW1, lc, F, R, hacc);
    // ---- L81 / L_Tr81 / L_Hg81 -> Comb[0:96]
    float comb[3][8];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        float in[1][8];
#pragma unroll
        for (int i = 0; i < 8; ++i) in[0][i] = R[b][i];
        float o[2][4];
        if (b == 0) dense<PREC, ML_TRUNK + 8>(net, in, o, lc.lane, true);
        else if (b == 1) dense<PREC, ML_TRUNK + 17>(net, in, o, lc.lane, true);
        else dense<PREC, ML_TRUNK + 26>(net, in, o, lc.lane, true);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            comb[b][i] = o[0][i];
            comb[b][4 + i] = o[1][i];
        }
    }
    // ---- DI stack: points 160..191 gathered at once; [160,168) lanes q<2 slots 0..3,
    // [168,176) lanes q>=2 slots 0..3, [176,184) lanes q<2 slots 4..7, [184,192) q>=2 4..7
    float D[3][8];
    src.get(lc, 160, true, D);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) R[b][i] = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) G[b][i] = i < 4 ? D[b][i] : 0.0f;
    di_block<PREC, 0>(net, sG, sW1, lc, G, R, hacc);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) G[b][i] = i < 4 ? xor32(D[b][i], lc.lane) : 0.0f;
    di_block<PREC, 1>(net, sG, sW1, lc, G, R, hacc);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) G[b][i] = i < 4 ? D[b][4 + i] : 0.0f;
    di_block<PREC, 2>(net, sG, sW1, lc, G, R, hacc);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int i = 0; i < 8; ++i) G[b][i] = i < 4 ? xor32(D[b][4 + i], lc.lane) : 0.0f;
    di_block<PREC, 3>(net, sG, sW1, lc, G, R, hacc);
    // ---- LD41 heads: tile A = [LD41 rows 0..7 | LD_Tr41 rows 8..15], tile B = [LD_Hg41 | LGGS]
    float tA[4], tB[4];
    {
        float in0[1][8], in1[1][8], in2[1][8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            in0[0][i] = R[0][i];
            in1[0][i] = R[1][i];
            in2[0][i] = R[2][i];
        }
        float o0[1][4], o1[1][4], o2[1][4];
        dense<PREC, ML_LD41A_VAL>(net, in0, o0, lc.lane, false);
        dense<PREC, ML_LD41A_TR>(net, in1, o1, lc.lane, false);
        dense<PREC, ML_LD41B_HG>(net, in2, o2, lc.lane, true);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            tA[i] = fmaxf(o0[0][i] + o1[0][i], 0.0f);
            tB[i] = o2[0][i];
        }
    }
    // ---- tail per distinct colour channel (radiancePredict.cu:411-421).  scatterrate is
    // uniform over the tile (one view): the first channel's tail is inlined, a second or third
    // distinct channel runs out of line
    float P0 = tail<PREC>(net, W, sW1, lc, lc.sp.x, hacc, comb, tA, tB);
    float P1 = P0, P2 = P0;
    if (lc.need_y) P1 = tail_call<PREC>(nt, W, sW1, lc, lc.sp.y, hacc, comb, tA, tB);
    if (lc.need_z) P2 = tail_call<PREC>(nt, W, sW1, lc, lc.sp.z, hacc, comb, tA, tB);
    if (lc.sr.y == lc.sr.x) P1 = P0;
    if (lc.sr.z == lc.sr.x) P2 = P0;
    else if (lc.sr.z == lc.sr.y) P2 = P1;
    v3 r;
    r.x = fmaxf(expf(P0) - 1.0f, 0.0f) * lc.sp.x;
    r.y = fmaxf(expf(P1) - 1.0f, 0.0f) * lc.sp.y;
    r.z = fmaxf(expf(P2) - 1.0f, 0.0f) * lc.sp.z;
    return r;
}

__device__ inline void load_shared(float* sW1, float* sG, StencilEntry* sst, const float* __restrict__ W,
                                   const StencilEntry* __restrict__ stencil) {
    for (int i = threadIdx.x; i < 1200; i += blockDim.x) sW1[i] = W[net::LSEFin1W + i];
    for (int i = threadIdx.x; i < kGates * kGateFloats; i += blockDim.x) {
        const int gi = i / kGateFloats, k = i % kGateFloats;
        sG[i] = k < 64 ? W[gate_w0(gi) + k] : W[gate_w1(gi) + k - 64];
    }
    if (sst)
        for (int i = threadIdx.x; i < kPC; i += blockDim.x) sst[i] = stencil[i];
    __syncthreads();
}

// Copies the variant's LDS weight image (built by mfma_net_create) into LDS.
template <int PREC>
__device__ inline void stage_weights(const MfmaNet& net, u32x4_t* sWf) {
    constexpr int n = lds_weight_bytes(PREC) / 16;
    const u32x4_t* src = (const u32x4_t*)net.img[base_prec(PREC)];
    for (int i = threadIdx.x; i < n; i += blockDim.x) sWf[i] = src[i];
}

// PREC: a precision id, plus kLdsW for the variant with LDS-resident weights (one 768-thread
// workgroup per CU); FAST: the fast gather of dev_gather_fast.hpp.
template <int PREC, bool FAST>
__global__ __launch_bounds__(lds_weights(PREC) ? 768 : 256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_mlp_mfma_active(MfmaNet net, const DevVolume* __restrict__ vols,
                                                         const ViewSlot* __restrict__ slots,
                                                         const StencilEntry* __restrict__ stencil,
                                                         const FastStencil* __restrict__ fast,
                                                         const float* __restrict__ lut_rows,
                                                         const float* __restrict__ W, const ActiveRec* __restrict__ recs,
                                                         const unsigned* __restrict__ counter) {
    __shared__ float sW1[1200];
    __shared__ float sG[kGates * kGateFloats];
    __shared__ StencilEntry sst[FAST ? 1 : kPC];
    __shared__ FastStencil sfs[FAST ? kPC : 1];
    __shared__ u32x4_t sWf[lds_weights(PREC) ? lds_weight_bytes(PREC) / 16 : 1];
    static_assert(!lds_weights(PREC) || lds_weight_bytes(PREC) + kLdsOther <= kLdsBudget,
                  "LDS budget of the LDS-weight kernel");
    if constexpr (lds_weights(PREC)) stage_weights<PREC>(net, sWf);
    if (FAST) {
        for (int i = threadIdx.x; i < kPC; i += blockDim.x) sfs[i] = fast[i];
        load_shared(sW1, sG, nullptr, W, nullptr);
    } else {
        load_shared(sW1, sG, sst, W, stencil);
    }
    const unsigned n = *counter;
    const int lane = threadIdx.x & 63;
    const int wave = xcd_swizzle(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const unsigned nwaves = gridDim.x * (blockDim.x >> 6);
    // full rounds: every wave one tile, an XCD's waves on a contiguous range of tiles (L2
    // locality).  The last, partial round takes the leftover tiles in an order that spreads them
    // over the workgroups first (one per CU, then a second ...), so it neither lands on the first
    // XCD alone nor queues behind busy waves of the same CU
    const unsigned ntiles = (n + 15u) / 16u, kfull = ntiles / nwaves;
    const unsigned tail = (threadIdx.x >> 6) * gridDim.x + blockIdx.x;
    for (unsigned k = 0;; ++k) {
        unsigned tile;
        if (k < kfull) {
            tile = k * nwaves + (unsigned)wave;
        } else {
            tile = kfull * nwaves + tail;
            if (k > kfull || tile >= ntiles) break;
        }
        LaneCtx lc;
        lc.lane = lane;
        lc.c = lane & 15;
        lc.q = lane >> 4;
        const unsigned idx = tile * 16u + lc.c;
        // all 16 records of a tile come from one view (padded compaction)
        const int view = __builtin_amdgcn_readfirstlane(recs[tile * 16u].view);
        const ActiveRec r = recs[idx];
        lc.ok = r.pixel >= 0;
        const ViewSlot& sl = slots[view];
        const ViewDev& vw = sl.vw;
        const ViewConst& vc = sl.vc;
        PixelState ps;
        ps.pos = mk3(r.pos[0], r.pos[1], r.pos[2]);
        ps.X = mk3(r.X[0], r.X[1], r.X[2]);
        ps.Y = mk3(r.Y[0], r.Y[1], r.Y[2]);
        ps.Z = mk3(r.Z[0], r.Z[1], r.Z[2]);
        ps.L = vc.L;
        ps.LX = vc.LX;
        ps.dist = r.dist;
        ps.active = 1;
        ps.taps = 0;
        lc.g = vw.g;
        lc.gamma = gamma_of(ps);
        lc.sr = vc.sr;
        lc.sp = vc.sp;
        lc.need_y = vc.need_y != 0;
        lc.need_z = vc.need_z != 0;
        MfmaNet nt = net;
        nt.bf16 = opaque(net.bf16);
        nt.bf16_lo = opaque(net.bf16_lo);
        nt.fp8 = opaque(net.fp8);
        nt.bias = opaque(net.bias);
        nt.lds = opaque_v(sWf);
        v3 rgb;
        if (FAST) {
            const FastGatherSrc src = fast_src(vols[sl.vol], vc, vw.g, ps, opaque_v(sfs), lut_rows);
            rgb = run_tile<PREC>(nt, opaque(W), opaque_v(sG), opaque_v(sW1), lc, src);
        } else {
            GatherSrc src;
            src.vol = &vols[sl.vol];
            src.st = sst;
            src.g = vw.g;
            src.alpha = vw.alpha;
            src.ps = ps;
            rgb = run_tile<PREC>(nt, opaque(W), opaque_v(sG), opaque_v(sW1), lc, src);
        }
        if (lc.ok && lc.q == 0) {
            v3 o = (rgb * vw.light_color) * vw.scatter_rate;
            sl.out[r.pixel] = make_float4(fmaxf(0.0f, o.x), fmaxf(0.0f, o.y), fmaxf(0.0f, o.z), fmaxf(0.0f, r.dist));
        }
    }
}

// The fast gather alone: for every active record its 577 features [X_Val | X_Val_Hg | X_Val_Sub |
// gamma] into out[pixel] of the record's view (parity tooling for the fused kernels' gather).
__global__ __launch_bounds__(256) void k_fast_features(const DevVolume* __restrict__ vols,
                                                       const ViewSlot* __restrict__ slots,
                                                       const FastStencil* __restrict__ fast,
                                                       const float* __restrict__ lut_rows,
                                                       const ActiveRec* __restrict__ recs,
                                                       const unsigned* __restrict__ counter, float* __restrict__ out) {
    __shared__ FastStencil sfs[kPC];
    for (int i = threadIdx.x; i < kPC; i += blockDim.x) sfs[i] = fast[i];
    __syncthreads();
    const unsigned n = *counter;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (unsigned tile = wave; tile * 16u < n; tile += nwaves) {
        LaneCtx lc;
        lc.lane = lane;
        lc.c = lane & 15;
        lc.q = lane >> 4;
        const ActiveRec r = recs[tile * 16u + lc.c];
        lc.ok = r.pixel >= 0;
        const ViewSlot& sl = slots[r.view];
        PixelState ps;
        ps.pos = mk3(r.pos[0], r.pos[1], r.pos[2]);
        ps.X = mk3(r.X[0], r.X[1], r.X[2]);
        ps.Y = mk3(r.Y[0], r.Y[1], r.Y[2]);
        ps.Z = mk3(r.Z[0], r.Z[1], r.Z[2]);
        ps.L = sl.vc.L;
        ps.LX = sl.vc.LX;
        const FastGatherSrc src = fast_src(vols[sl.vol], sl.vc, sl.vw.g, ps, sfs, lut_rows);
        float* row = out + (size_t)(lc.ok ? r.pixel : 0) * kNFEAT;
        constexpr int p0s[8] = {0, 16, 32, 48, 64, 96, 128, 160};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool wide = p0s[j] >= 64;
            float F[3][8];
            src.get(lc, p0s[j], wide, F);
            if (!lc.ok) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (h == 1 && !wide) break;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int k = p0s[j] + 16 * h + 4 * lc.q + i;
                    row[k] = F[0][4 * h + i];
                    row[kPC + k] = F[2][4 * h + i];
                    row[2 * kPC + k] = F[1][4 * h + i];
                }
            }
        }
        if (lc.ok && lc.q == 0) row[3 * kPC] = gamma_of(ps);
    }
}

hipError_t launch_fast_features(const DevVolume* vols, const ViewSlot* slots, const void* fast_stencil,
                                const float* lut_rows, const ActiveRec* recs, const unsigned* counter, int grid,
                                float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fast_features, dim3(grid), dim3(256), 0, s, vols, slots, (const FastStencil*)fast_stencil,
                       lut_rows, recs, counter, out);
    return hipGetLastError();
}

template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_mlp_mfma_features(
    MfmaNet net, const float* __restrict__ W, const float* __restrict__ feats, int n, v3 sr, float g,
    float* __restrict__ rgb) {
    __shared__ float sW1[1200];
    __shared__ float sG[kGates * kGateFloats];
    load_shared(sW1, sG, nullptr, W, nullptr);
    const v3 sp = mk3(powf(sr.x, 4.0f), powf(sr.y, 4.0f), powf(sr.z, 4.0f));
    const int lane = threadIdx.x & 63;
    const int wave = xcd_swizzle(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (int tile = wave; tile * 16 < n; tile += nwaves) {
        LaneCtx lc;
        lc.lane = lane;
        lc.c = lane & 15;
        lc.q = lane >> 4;
        const int idx = tile * 16 + lc.c;
        lc.ok = idx < n;
        RowSrc src;
        src.row = feats + (size_t)(lc.ok ? idx : 0) * kNFEAT;
        lc.g = g;
        lc.gamma = lc.ok ? src.row[3 * kPC] : 0.0f;
        lc.sr = sr;
        lc.sp = sp;
        lc.need_y = sr.y != sr.x;
        lc.need_z = sr.z != sr.x && sr.z != sr.y;
        MfmaNet nt = net;
        nt.bf16 = opaque(net.bf16);
        nt.bf16_lo = opaque(net.bf16_lo);
        nt.fp8 = opaque(net.fp8);
        nt.bias = opaque(net.bias);
        v3 o = run_tile<PREC>(nt, opaque(W), opaque_v(sG), opaque_v(sW1), lc, src);
        if (!calibrating(PREC) && lc.ok && lc.q == 0) {
            rgb[3 * (size_t)idx] = o.x;
            rgb[3 * (size_t)idx + 1] = o.y;
            rgb[3 * (size_t)idx + 2] = o.z;
        }
    }
}

// grid: workgroups of 256 threads (L2 weights); the LDS-weight variants take `grid` as the
// number of 768-thread workgroups (one per CU).
template <int PREC>
static void launch_active_prec(bool fast, const MfmaNet& net, const DevVolume* vols, const ViewSlot* slots,
                               const StencilEntry* stencil, const FastStencil* fst, const float* rows,
                               const float* W, const ActiveRec* recs, const unsigned* counter, int grid,
                               hipStream_t s) {
    const int block = lds_weights(PREC) ? 768 : 256;
    if (fast)
        hipLaunchKernelGGL((k_mlp_mfma_active<PREC, true>), dim3(grid), dim3(block), 0, s, net, vols, slots, stencil,
                           fst, rows, W, recs, counter);
    else
        hipLaunchKernelGGL((k_mlp_mfma_active<PREC, false>), dim3(grid), dim3(block), 0, s, net, vols, slots, stencil,
                           fst, rows, W, recs, counter);
}

hipError_t launch_mlp_mfma_active(int precision, bool fast, bool lds_w, const MfmaNet& net, const DevVolume* vols,
                                  const ViewSlot* slots, const StencilEntry* stencil, const void* fast_stencil,
                                  const float* lut_rows, const float* W, const ActiveRec* recs,
                                  const unsigned* counter, int grid, hipStream_t s) {
    const FastStencil* fst = (const FastStencil*)fast_stencil;
#define MRP_LAUNCH(P) launch_active_prec<P>(fast, net, vols, slots, stencil, fst, lut_rows, W, recs, counter, grid, s)
    if (precision == MRP_BF16) {
        if (lds_w) MRP_LAUNCH(MRP_BF16 | kLdsW);
        else MRP_LAUNCH(MRP_BF16);
    } else if (precision == MRP_BF16_1T) {
        if (lds_w) MRP_LAUNCH(MRP_BF16_1T | kLdsW);
        else MRP_LAUNCH(MRP_BF16_1T);
    } else {
        if (lds_w) MRP_LAUNCH(MRP_FP8 | kLdsW);
        else MRP_LAUNCH(MRP_FP8);
    }
#undef MRP_LAUNCH
    return hipGetLastError();
}

hipError_t launch_mlp_mfma_features(int precision, const MfmaNet& net, const float* W, const float* feats, int n,
                                    const float sr[3], float g, float* rgb, hipStream_t s) {
    int tiles = (n + 15) / 16;
    int grid = (tiles + 3) / 4;
    if (grid > 4096) grid = 4096;
    if (precision == MRP_BF16)
        hipLaunchKernelGGL(k_mlp_mfma_features<MRP_BF16>, dim3(grid), dim3(256), 0, s, net, W, feats, n,
                           mk3(sr[0], sr[1], sr[2]), g, rgb);
    else if (precision == MRP_BF16_1T)
        hipLaunchKernelGGL(k_mlp_mfma_features<MRP_BF16_1T>, dim3(grid), dim3(256), 0, s, net, W, feats, n,
                           mk3(sr[0], sr[1], sr[2]), g, rgb);
    else
        hipLaunchKernelGGL(k_mlp_mfma_features<MRP_FP8>, dim3(grid), dim3(256), 0, s, net, W, feats, n,
                           mk3(sr[0], sr[1], sr[2]), g, rgb);
    return hipGetLastError();
}

hipError_t launch_mlp_mfma_calibrate(const MfmaNet& net, const float* W, const float* feats, int n, const float sr[3],
                                     float g, hipStream_t s) {
    int tiles = (n + 15) / 16;
    int grid = (tiles + 3) / 4;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_mlp_mfma_features<MRP_BF16 | kCal>, dim3(grid), dim3(256), 0, s, net, W, feats, n,
                       mk3(sr[0], sr[1], sr[2]), g, (float*)nullptr);
    return hipGetLastError();
}

}  // namespace mrp
