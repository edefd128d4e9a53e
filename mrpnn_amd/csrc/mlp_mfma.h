// MFMA form of the MRPNN network (BF16 and FP8 e4m3, FP32 accumulation) on gfx950.
//
// Tile = 16 pixels per wavefront; lane = c + 16*q with c the pixel column and q in 0..3.
// Every dense layer is v_mfma_f32_16x16x32_{bf16,fp8} with weights as A (16 output rows per
// M-tile) and activations as B (16 pixels).  Activation vectors live in registers in the
// D layout of the producing layer: M-tile t contributes rows 16t+4q+i (i = 0..3) to lane
// (c, q).  The K slots of a 32-wide chunk k are ordered so that slot 8q+i <-> input row
// 32k+4q+i and slot 8q+4+i <-> row 32k+16+4q+i: exactly the rows a lane already holds, so
// D feeds the next layer's B without any cross-lane movement.  Layers with K = 16 (or 8)
// use one chunk whose slots 8q+4..7 are zero.  The host packs each weight fragment (64
// lanes x 8 values) in that order (pack in mlp_mfma.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fused.h"

namespace mrp {

enum MfmaLayerId {
    ML_TRUNK = 0,         // + 9*branch + k, k = 0..8 : L01 L11 L21 L31 L41 L51 L61 L71 L81
    ML_DI = 27,           // + 4*branch + k, k = 0..3 : LD01 LD11 LD21 LD31
    ML_LD41A_VAL = 39,    // LD41 into rows 0..7 of tile A
    ML_LD41A_TR = 40,     // LD_Tr41 into rows 8..15 of tile A
    ML_LD41B_HG = 41,     // LD_Hg41 into rows 0..7 of tile B
    ML_FIN2 = 42,         // LSEFin2 (6x16), no bias
    ML_LC0 = 43,
    ML_LC1 = 44,
    ML_LC2 = 45,
    ML_LX = 46,
    ML_LXK = 47,          // + k, k = 0..11
    ML_LX12 = 59,
    ML_COUNT = 60
};

// Kernel variants: a precision id (MRP_BF16, MRP_FP8, MRP_BF16_1T) plus kLdsW when the weight
// fragments are staged in LDS (one 768-thread workgroup per CU shares one copy) instead of
// being read through the L1/L2 path for every tile.
constexpr int kLdsW = 16;
// kCal: the BF16 network that also records, per layer, the largest |input| it quantises
// (the calibration of the FP8 activation scales).
constexpr int kCal = 32;
constexpr int base_prec(int v) { return v & (kLdsW - 1); }
constexpr bool lds_weights(int v) { return (v & kLdsW) != 0; }
constexpr bool calibrating(int v) { return (v & kCal) != 0; }

// Compile-time shape of every layer: rows, inputs, first destination row inside the padded
// 16-row M-tiles, and whether K comes in 32-wide chunks (else K <= 16, one chunk whose
// slots 8q+4..7 are zero).
struct MfmaShape {
    int M, K, roff, k32;
};
constexpr MfmaShape mfma_shape(int id) {
    if (id < ML_DI) {
        const int k = (id - ML_TRUNK) % 9;
        const int n = k < 2 ? 8 : (k < 5 ? 16 : 32);
        return {n, n, 0, n == 32};
    }
    if (id < ML_LD41A_VAL) return {8, 8, 0, 0};
    if (id == ML_LD41A_VAL) return {8, 8, 0, 0};
    if (id == ML_LD41A_TR) return {8, 8, 8, 0};
    if (id == ML_LD41B_HG) return {8, 8, 0, 0};
    if (id == ML_FIN2) return {6, 16, 0, 0};
    if (id == ML_LC0) return {128, 128, 0, 1};
    if (id == ML_LC1) return {64, 128, 0, 1};
    if (id == ML_LC2) return {32, 64, 0, 1};
    if (id == ML_LX) return {16, 32, 0, 1};
    if (id < ML_LX12) return {16, 16, 0, 0};
    return {1, 16, 0, 0};
}
constexpr int mfma_nm(int id) { return (mfma_shape(id).roff + mfma_shape(id).M + 15) / 16; }
constexpr int mfma_nk(int id) { return mfma_shape(id).k32 ? mfma_shape(id).K / 32 : 1; }
constexpr bool mfma_has_bias(int id) { return id != ML_FIN2; }
// first fragment (64 lanes x 8 values) and bias offset (16 floats per M-tile) of a layer
constexpr int mfma_frag(int id) {
    int f = 0;
    for (int i = 0; i < id; ++i) f += mfma_nm(i) * mfma_nk(i);
    return f;
}
constexpr int mfma_bias(int id) {
    int b = 0;
    for (int i = 0; i < id; ++i)
        if (mfma_has_bias(i)) b += 16 * mfma_nm(i);
    return b;
}
constexpr int kMfmaFrags = mfma_frag(ML_COUNT);
constexpr int kMfmaBias = mfma_bias(ML_COUNT);

// MRP_BF16 adds a second bf16 weight term (lo = bf16(W - bf16(W))) on the tail layers after
// LC0 (LC1, LC2, LX, the LX chain); emulating the network on fBm-cloud features, these carry
// most of what the second term buys: mean relative error 1.28 % single-term, 0.69 % with the
// whole tail two-term, 0.90 % with this set, and LC0's 32 fragments of lo terms stay out of LDS.
constexpr bool two_term(int v, int id) { return base_prec(v) == 1 && id >= ML_LC1; }

// LDS image of a variant (built on the host, copied verbatim by the kernel): the hi fragments
// of every layer, then the lo fragments of the two-term layers, then all biases.  A fragment
// is 64 lanes x 8 values (1 KB bf16, 512 B fp8); for BF16 the fragments of K <= 16 layers
// keep only slots 0..3 per lane (slots 4..7 are zero): 512 B.
constexpr bool compact_frag(int v, int id) { return base_prec(v) == 1 && !mfma_shape(id).k32; }
constexpr int frag_bytes(int v, int id) { return base_prec(v) == 2 ? 512 : (compact_frag(v, id) ? 512 : 1024); }
constexpr int lds_hi_off(int v, int id) {
    int o = 0;
    for (int i = 0; i < id; ++i) o += mfma_nm(i) * mfma_nk(i) * frag_bytes(v, i);
    return o;
}
constexpr int lds_lo_off(int v, int id) {
    int o = lds_hi_off(v, ML_COUNT);
    for (int i = 0; i < id; ++i)
        if (two_term(v, i)) o += mfma_nm(i) * mfma_nk(i) * frag_bytes(v, i);
    return o;
}
constexpr int lds_frag_bytes(int v) { return lds_weights(v) ? lds_lo_off(v, ML_COUNT) : 0; }
constexpr int lds_weight_bytes(int v) { return lds_weights(v) ? lds_frag_bytes(v) + kMfmaBias * 4 : 0; }
// the rest of the kernel's LDS (LSEFin1 and gate weights, stencil table) and the workgroup
// budget: a workgroup asking for 163,392 of the CU's 163,840 bytes measured at about half
// speed, one with 161,600 bytes at full speed
constexpr int kLdsOther = 24384, kLdsBudget = 161792;
struct MfmaNet {
    const void* lds = nullptr;       // device: the workgroup's LDS copy of the image (set in the kernel)
    const uint4* img[4] = {nullptr, nullptr, nullptr, nullptr};  // LDS images by precision id
    const uint4* bf16 = nullptr;     // fragments, 16 bytes per lane: bf16(W)
    const uint4* bf16_lo = nullptr;  // bf16(W - bf16(W)): second term of the two-term weights
    const uint2* fp8 = nullptr;   // fragments, 8 bytes per lane
    const float* bias = nullptr;
    // FP8 per layer: weights stored as fp8(W * 2^wexp), inputs as fp8(clamp(x) * 2^aexp);
    // f8[id] = {2^-aexp (the divisor of the scaled conversion), 448 * 2^-aexp (input clamp),
    // 2^-(aexp + wexp) (product scale), 0}.  The same allocation continues with ML_COUNT
    // unsigned words: the per-layer max |input| (float bits) that the kCal kernels record.
    // (The struct is a kernel argument copied per tile: host-side state lives elsewhere.)
    const float4* f8 = nullptr;
};
__host__ __device__ inline unsigned* mfma_cal_words(const MfmaNet& n) { return (unsigned*)(n.f8 + ML_COUNT); }

// Packs the weights; wexp receives the per-layer FP8 weight exponents.
int mfma_net_create(const float* weights_fp32, MfmaNet* out, int* wexp);
void mfma_net_destroy(MfmaNet* net);
// Installs FP8 activation exponents (aexp[id]: layer id's inputs are stored as x * 2^aexp[id]).
// Blocking; the caller orders it after work that reads the table.
int mfma_net_set_fp8_exps(const MfmaNet& net, const int* wexp, const int* aexp);
// The exponent max-abs scaling gives a layer whose inputs reach amax: the largest e with
// amax * 2^e <= 448 (e4m3's largest finite value); 0 for amax = 0.
int fp8_exp_for_max(float amax);

// Host encoders used by the packer (exposed for tests through the C ABI of the library).
uint16_t host_f32_to_bf16(float f);
uint8_t host_f32_to_e4m3(float f);

hipError_t launch_mlp_mfma_active(int precision, bool fast, bool lds_w, const MfmaNet& net, const DevVolume* vols,
                                  const ViewSlot* slots, const StencilEntry* stencil, const void* fast_stencil,
                                  const float* lut_rows, const float* W, const ActiveRec* recs,
                                  const unsigned* counter, int grid, hipStream_t s);
// The fast gather's 577 features of every active record into out (row = record pixel).
hipError_t launch_fast_features(const DevVolume* vols, const ViewSlot* slots, const void* fast_stencil,
                                const float* lut_rows, const ActiveRec* recs, const unsigned* counter, int grid,
                                float* out, hipStream_t s);
hipError_t launch_mlp_mfma_features(int precision, const MfmaNet& net, const float* W, const float* feats, int n,
                                    const float sr[3], float g, float* rgb, hipStream_t s);
// The BF16 network on n feature rows recording per layer the largest |input| into net.cal
// (no output is written).
hipError_t launch_mlp_mfma_calibrate(const MfmaNet& net, const float* W, const float* feats, int n, const float sr[3],
                                     float g, hipStream_t s);

}  // namespace mrp
