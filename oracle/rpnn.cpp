// Oracle of the RPNN comparison predictor (TEST INFRASTRUCTURE ONLY):
// RadiancePredict_RPNN (core/radiancePredict.cu:567-722) called from NNPredict<RPNN>
// (core/render.cuh:808-875), weights of core/NNWeight_RPNN.cuh (the rpnn_weights.bin asset).
//   descriptor: 10 levels x 225 density samples on a 5 x 5 x 9 grid in the light frame
//               (eX = cross(LX, X), eY = cross(eX, LX), LX), spacing 2^(l-12), mip
//               max(min(l - 2, 9), 0), value density * alpha / 64;       radiancePredict.cu:580-611
//   network:    per level B = V_l [226] (+ W_l C + b), ReLU, C = (C +) U_l B + b, ReLU; then
//               F0, F1 (ReLU) and F2; out = max(exp(y) - 1, 0);           radiancePredict.cu:624-721
// Operators in the reference's order: each output sums its products from input 0 upwards
// starting at 0 and adds the bias last (Linear / Linear_Unbias, radiancePredict.cu:88-117).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "mrp_oracle.hpp"

using namespace oracle;

namespace {
// exp of the FP32 network: exp_f32 (fexp.h, the device's operations); of the FP64 network: libm
template <class T> inline T exp_of(T x);
template <> inline float exp_of<float>(float x) { return mrp::exp_f32(x); }
template <> inline double exp_of<double>(double x) { return exp(x); }


constexpr int L = 10, S = 225, NIN = 226, B = 200;
std::vector<float> g_rw;  // the asset in header order
bool g_ok = false;

struct Layout {
    const float* V[L];
    const float* W[L];
    const float* Wb[L];
    const float* U[L];
    const float* Ub[L];
    const float *F0, *F0b, *F1, *F1b, *F2, *F2b;
};
Layout layout(const float* w) {
    Layout o{};
    size_t off = 0;
    for (int l = 0; l < L; ++l, off += (size_t)B * NIN) o.V[l] = w + off;
    for (int l = 1; l < L; ++l) {
        o.W[l] = w + off;
        off += (size_t)B * B;
        o.Wb[l] = w + off;
        off += B;
    }
    for (int l = 0; l < L; ++l) {
        o.U[l] = w + off;
        off += (size_t)B * B;
        o.Ub[l] = w + off;
        off += B;
    }
    o.F0 = w + off;
    off += (size_t)B * B;
    o.F0b = w + off;
    off += B;
    o.F1 = w + off;
    off += (size_t)B * B;
    o.F1b = w + off;
    off += B;
    o.F2 = w + off;
    off += B;
    o.F2b = w + off;
    return o;
}

template <class T>
void linear(const T* in, T* out, int from, int to, const float* w, const float* bias) {
    for (int i = 0; i < to; ++i) {
        T acc = (T)0;
        for (int j = 0; j < from; ++j) acc += in[j] * (T)w[j + i * from];
        if (bias) acc += (T)bias[i];
        out[i] = acc;
    }
}
template <class T>
void relu(T* x, int n) {
    for (int i = 0; i < n; ++i) x[i] = x[i] >= (T)0 ? x[i] : (T)0;
}

template <class T>
T rpnn_net(const Layout& w, const float* desc /* 2250 */, float gamma) {
    T A[NIN], Bv[B], C[B], C2[B];
    A[S] = (T)gamma;
    for (int l = 0; l < L; ++l) {
        for (int k = 0; k < S; ++k) A[k] = (T)desc[l * S + k];  // Rep
        linear<T>(A, Bv, NIN, B, w.V[l], nullptr);
        if (l > 0) {
            linear<T>(C, C2, B, B, w.W[l], w.Wb[l]);
            for (int i = 0; i < B; ++i) Bv[i] = Bv[i] + C2[i];
        }
        relu<T>(Bv, B);
        if (l == 0) {
            linear<T>(Bv, C, B, B, w.U[0], w.Ub[0]);
        } else {
            linear<T>(Bv, C2, B, B, w.U[l], w.Ub[l]);
            for (int i = 0; i < B; ++i) C[i] = C[i] + C2[i];
        }
        relu<T>(C, B);
    }
    linear<T>(C, C2, B, B, w.F0, w.F0b);
    relu<T>(C2, B);
    linear<T>(C2, C, B, B, w.F1, w.F1b);
    relu<T>(C, B);
    T y;
    linear<T>(C, &y, B, 1, w.F2, w.F2b);
    relu<T>(&y, 1);  // LinearReLU for F2 as well (radiancePredict.cu:716)
    T o = exp_of<T>(y) - (T)1;
    return o > (T)0 ? o : (T)0;
}

}  // namespace

namespace oracle {
// descriptor of one shading point: 2250 values, gamma = acos(dot(X, LX)) (not clamped)
void rpnn_descriptor(const Volume& v, f3 pos, f3 X, f3 LX, float alpha, float* desc, float* gamma) {
    *gamma = acosf(dot(X, LX));
    const f3 eX = normalize(cross(LX, X));
    const f3 eY = normalize(cross(eX, LX));
    const float desc_size0 = 0.5f / 1024.0f;
    float scale = 0.5f * desc_size0;
    float mip_level = 0.0f;
    for (int i = 0; i < L; ++i) {
        int k = 0;
        const float cur = fmaxf(fminf(mip_level - 2.0f, 9.0f), 0.0f);
        const int mip = (int)(cur + 0.001f);
        for (int z = -2; z <= 6; ++z)
            for (int y = -2; y <= 2; ++y)
                for (int x = -2; x <= 2; ++x) {
                    const f3 offset = (eX * (float)x + eY * (float)y + LX * (float)z) * scale;
                    const f3 p = pos + offset;
                    desc[i * S + k++] = mip_density(v, mip, p) * alpha / 64.0f;
                }
        scale *= 2.0f;
        mip_level += 1.0f;
    }
}

bool rpnn_ready() { return g_ok; }
float rpnn_predict_f32(const float* desc, float gamma) { return rpnn_net<float>(layout(g_rw.data()), desc, gamma); }
}  // namespace oracle

extern "C" {

int orc_rpnn_load(const float* w, int total) {
    if (total != 1296401) return 1;
    g_rw.assign(w, w + total);
    g_ok = true;
    return 0;
}

// n rows of 2251 floats (2250 descriptor values, gamma) -> n outputs
int orc_rpnn_predict_f32(const float* rows, int n, float* out) {
    if (!g_ok) return 1;
    Layout w = layout(g_rw.data());
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) out[i] = rpnn_net<float>(w, rows + (size_t)i * 2251, rows[(size_t)i * 2251 + 2250]);
    return 0;
}

int orc_rpnn_predict_f64(const float* rows, int n, double* out) {
    if (!g_ok) return 1;
    Layout w = layout(g_rw.data());
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) out[i] = rpnn_net<double>(w, rows + (size_t)i * 2251, rows[(size_t)i * 2251 + 2250]);
    return 0;
}

}
