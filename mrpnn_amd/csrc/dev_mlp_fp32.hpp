// Scalar FP32 MRPNN network, RadiancePredictInternal (core/radiancePredict.cu:301-422),
// one pixel per thread, in the reference's structure and summation order:
//   * LinearReLU_SFL / LinearReLU<Size>: bias first, inputs ascending (:129-224);
//   * untemplated LinearReLU / *_Unbias / LinearSm_Unbias: zero start, bias last (:41-86);
//   * every multiply-accumulate is one fmaf (nvcc contracts `Out += In * W`);
//   * AVGMAX: sequential fp32 sum divided by n, max from -10000 (:21-33);
//   * sigmoid as 1/(1+exp(-x)) (:84), exp(P) of the output and scatterrate^4 in single precision,
//     exp from fexp.h (no double-precision arithmetic).
// Weights are the fp32 asset in header order, addressed by the offsets in net_layout.h;
// their addresses are wavefront-uniform, so loads go through the scalar cache.
// This is the drop-in baseline and the FP32 parity target; the fast path lives in the
// fused MFMA kernels.
#pragma once
#include "dev_common.hpp"
#include "fexp.h"
#include "net_layout.h"

namespace mrp {

struct Fp32Mlp {
    const float* __restrict__ W;
    const float* X;
    const float* S;
    const float* H;
    float A[160], B[160], C[160], A2[160], B2[160], C2[160];
    float pool[75], aw[16], awp[6], tap[8], comb[128];

    __device__ static float relu(float x) { return x >= 0.0f ? x : 0.0f; }
    // exp is exp_f32 (fexp.h), the same float operations as the CPU restatement: device and host
    // libm expf differ by an ulp, which exp(P)-1 amplifies past 1e-5 for small outputs.
    __device__ static float sigm(float x) { return 1.0f / (1.0f + exp_f32(-x)); }
    __device__ static float expr(float x) { return exp_f32(x); }

    __device__ void lin_relu_bf(const float* in, float* out, int from, int to, int w, int b) {
        for (int i = 0; i < to; ++i) {
            float acc = W[b + i];
            for (int j = 0; j < from; ++j) acc = fmaf(in[j], W[w + j + i * from], acc);
            out[i] = relu(acc);
        }
    }
    __device__ void lin_relu_bl(const float* in, float* out, int from, int to, int w, int b) {
        for (int i = 0; i < to; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < from; ++j) acc = fmaf(in[j], W[w + j + i * from], acc);
            acc += W[b + i];
            out[i] = relu(acc);
        }
    }
    __device__ void lin_relu_nb(const float* in, float* out, int from, int to, int w) {
        for (int i = 0; i < to; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < from; ++j) acc = fmaf(in[j], W[w + j + i * from], acc);
            out[i] = relu(acc);
        }
    }
    __device__ void lin_sigm_nb(const float* in, float* out, int from, int to, int w) {
        for (int i = 0; i < to; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < from; ++j) acc = fmaf(in[j], W[w + j + i * from], acc);
            out[i] = sigm(acc);
        }
    }
    __device__ void avgmax(const float* in, int n, int from, int w1, int w2) {
        float s = 0.0f, m = -10000.0f;
        for (int i = 0; i < n; ++i) {
            s += in[from + i];
            m = fmaxf(m, in[from + i]);
        }
        pool[w1] = s / (float)n;
        pool[w2] = m;
    }
    // SE (last < 0) / SERES (core/radiancePredict.cu:240-279)
    __device__ void block(int last, int from, int to, int n, int pid, int sew0, int sew1, int wA, int bA, int wB,
                          int bB, int wC, int bC) {
        avgmax(X, n, from, pid + 0, 36 + pid + 0);
        avgmax(S, n, from, pid + 1, 36 + pid + 1);
        avgmax(H, n, from, pid + 2, 36 + pid + 2);
        for (int i = 0; i < 3; ++i) tap[i] = pool[pid + i];
        for (int i = 0; i < 3; ++i) tap[3 + i] = pool[36 + pid + i];
        lin_relu_nb(tap, aw, 8, 8, sew0);
        lin_sigm_nb(aw, awp, 8, 3, sew1);
        for (int i = 0; i < n; ++i) {
            float a = X[from + i] * awp[0], b = S[from + i] * awp[1], c = H[from + i] * awp[2];
            if (last >= 0) { a = a + A2[last + i]; b = b + B2[last + i]; c = c + C2[last + i]; }
            A[to + i] = a;
            B[to + i] = b;
            C[to + i] = c;
        }
        lin_relu_bf(A + to, A2 + to, n, n, wA, bA);
        lin_relu_bf(B + to, B2 + to, n, n, wB, bB);
        lin_relu_bf(C + to, C2 + to, n, n, wC, bC);
    }
    __device__ void add3(int from, int to, int n) {  // ADD3 (:281-285)
        for (int i = 0; i < n; ++i) {
            A2[to + i] = A2[to + i] + X[from + i];
            B2[to + i] = B2[to + i] + S[from + i];
            C2[to + i] = C2[to + i] + H[from + i];
        }
    }
    __device__ void rep3(int from, int to, int n) {  // REP3 (:287-294)
        for (int i = 0; i < n; ++i) { A2[from + i] = A2[to + i]; B2[from + i] = B2[to + i]; C2[from + i] = C2[to + i]; }
        for (int i = 0; i < n; ++i) { A2[to + i] = X[to + i]; B2[to + i] = S[to + i]; C2[to + i] = H[to + i]; }
    }
    // SC(p) (:375-409)
    __device__ float tail(float srp_p) {
        using namespace net;
        pool[74] = srp_p;
        for (int i = 0; i < 120; ++i) A[i] = comb[i];
        lin_relu_bl(pool + 72, A + 120, 3, 8, LGGSW, LGGSB);
        lin_relu_nb(pool, aw, 75, 16, LSEFin1W);
        lin_sigm_nb(aw, awp, 16, 6, LSEFin2W);
        for (int i = 0; i < 32; ++i) A[i] *= awp[0];
        for (int i = 0; i < 32; ++i) A[32 + i] *= awp[1];
        for (int i = 0; i < 32; ++i) A[64 + i] *= awp[2];
        for (int i = 0; i < 8; ++i) A[96 + i] *= awp[3];
        for (int i = 0; i < 8; ++i) A[104 + i] *= awp[4];
        for (int i = 0; i < 8; ++i) A[112 + i] *= awp[5];
        lin_relu_bf(A, B, 128, 128, LC0W, LC0B);
        lin_relu_bf(B, A, 128, 64, LC1W, LC1B);
        lin_relu_bf(A, B, 64, 32, LC2W, LC2B);
        lin_relu_bf(B, A, 32, 16, LXW, LXB);
        const int lw[12] = {LX0W, LX1W, LX2W, LX3W, LX4W, LX5W, LX6W, LX7W, LX8W, LX9W, LX10W, LX11W};
        const int lb[12] = {LX0B, LX1B, LX2B, LX3B, LX4B, LX5B, LX6B, LX7B, LX8B, LX9B, LX10B, LX11B};
        for (int k = 0; k < 6; ++k) {
            lin_relu_bf(A, B, 16, 16, lw[2 * k], lb[2 * k]);
            lin_relu_bf(B, C, 16, 16, lw[2 * k + 1], lb[2 * k + 1]);
            for (int i = 0; i < 16; ++i) A[i] = A[i] + C[i];
        }
        lin_relu_bl(A, B, 16, 1, LX12W, LX12B);
        return B[0];
    }

    __device__ v3 run(float gamma, float g, v3 sr) {
        using namespace net;
        const int DI = 160, P_DI = 24;
        tap[6] = g;
        tap[7] = gamma;
        float srp[3] = {(sr.x * sr.x) * (sr.x * sr.x), (sr.y * sr.y) * (sr.y * sr.y), (sr.z * sr.z) * (sr.z * sr.z)};
        pool[72] = g;
        pool[73] = gamma;
        block(-1, 0, 0, 8, 0, LSE01W, LSE02W, L01W, L01B, L_Tr01W, L_Tr01B, L_Hg01W, L_Hg01B);
        add3(0, 0, 8);
        block(0, 8, 8, 8, 3, LSE11W, LSE12W, L11W, L11B, L_Tr11W, L_Tr11B, L_Hg11W, L_Hg11B);
        rep3(0, 8, 8);
        block(0, 16, 16, 16, 6, LSE21W, LSE22W, L21W, L21B, L_Tr21W, L_Tr21B, L_Hg21W, L_Hg21B);
        add3(16, 16, 16);
        block(16, 32, 32, 16, 9, LSE31W, LSE32W, L31W, L31B, L_Tr31W, L_Tr31B, L_Hg31W, L_Hg31B);
        add3(32, 32, 16);
        block(32, 48, 48, 16, 12, LSE41W, LSE42W, L41W, L41B, L_Tr41W, L_Tr41B, L_Hg41W, L_Hg41B);
        rep3(32, 48, 16);
        block(32, 64, 64, 32, 15, LSE51W, LSE52W, L51W, L51B, L_Tr51W, L_Tr51B, L_Hg51W, L_Hg51B);
        add3(64, 64, 32);
        block(64, 96, 96, 32, 18, LSE61W, LSE62W, L61W, L61B, L_Tr61W, L_Tr61B, L_Hg61W, L_Hg61B);
        add3(96, 96, 32);
        block(96, 128, 128, 32, 21, LSE71W, LSE72W, L71W, L71B, L_Tr71W, L_Tr71B, L_Hg71W, L_Hg71B);
        add3(128, 128, 32);
        lin_relu_bf(A2 + 128, comb + 0, 32, 32, L81W, L81B);
        lin_relu_bf(B2 + 128, comb + 32, 32, 32, L_Tr81W, L_Tr81B);
        lin_relu_bf(C2 + 128, comb + 64, 32, 32, L_Hg81W, L_Hg81B);
        block(-1, DI + 0, 0, 8, P_DI + 0, LDSE01W, LDSE02W, LD01W, LD01B, LD_Tr01W, LD_Tr01B, LD_Hg01W, LD_Hg01B);
        add3(DI + 0, 0, 8);
        block(0, DI + 8, 8, 8, P_DI + 3, LDSE11W, LDSE12W, LD11W, LD11B, LD_Tr11W, LD_Tr11B, LD_Hg11W, LD_Hg11B);
        add3(DI + 8, 8, 8);
        block(8, DI + 16, 16, 8, P_DI + 6, LDSE21W, LDSE22W, LD21W, LD21B, LD_Tr21W, LD_Tr21B, LD_Hg21W, LD_Hg21B);
        add3(DI + 16, 16, 8);
        block(16, DI + 24, 24, 8, P_DI + 9, LDSE31W, LDSE32W, LD31W, LD31B, LD_Tr31W, LD_Tr31B, LD_Hg31W, LD_Hg31B);
        add3(DI + 24, 24, 8);
        lin_relu_bf(A2 + 24, comb + 96, 8, 8, LD41W, LD41B);
        lin_relu_bf(B2 + 24, comb + 104, 8, 8, LD_Tr41W, LD_Tr41B);
        lin_relu_bf(C2 + 24, comb + 112, 8, 8, LD_Hg41W, LD_Hg41B);
        float P0 = tail(srp[0]);
        float P1 = (sr.y == sr.x) ? P0 : tail(srp[1]);
        float P2 = (sr.z == sr.x) ? P0 : ((sr.z == sr.y) ? P1 : tail(srp[2]));
        return mk3(fmaxf(expr(P0) - 1.0f, 0.0f) * srp[0], fmaxf(expr(P1) - 1.0f, 0.0f) * srp[1],
                   fmaxf(expr(P2) - 1.0f, 0.0f) * srp[2]);
    }
};

}  // namespace mrp
