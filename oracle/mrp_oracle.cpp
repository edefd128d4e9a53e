// CPU oracle implementation.  TEST INFRASTRUCTURE ONLY (see mrp_oracle.hpp).
#include "mrp_oracle.hpp"

#include <float.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <string>
#include <cmath>

namespace oracle {

// Array order of core/NNWeight.cuh (the weight asset keeps this order).
static const char* kWeightNames[146] = {
    "LD01W", "LD01B", "LD11W", "LD11B", "LD21W", "LD21B", "LD31W", "LD31B", "LD41W", "LD41B",
    "LD_Tr01W", "LD_Tr01B", "LD_Tr11W", "LD_Tr11B", "LD_Tr21W", "LD_Tr21B", "LD_Tr31W", "LD_Tr31B",
    "LD_Tr41W", "LD_Tr41B", "LD_Hg01W", "LD_Hg01B", "LD_Hg11W", "LD_Hg11B", "LD_Hg21W", "LD_Hg21B",
    "LD_Hg31W", "LD_Hg31B", "LD_Hg41W", "LD_Hg41B", "L01W", "L01B", "L11W", "L11B", "L21W", "L21B",
    "L31W", "L31B", "L41W", "L41B", "L51W", "L51B", "L61W", "L61B", "L71W", "L71B", "L81W", "L81B",
    "L_Tr01W", "L_Tr01B", "L_Tr11W", "L_Tr11B", "L_Tr21W", "L_Tr21B", "L_Tr31W", "L_Tr31B", "L_Tr41W",
    "L_Tr41B", "L_Tr51W", "L_Tr51B", "L_Tr61W", "L_Tr61B", "L_Tr71W", "L_Tr71B", "L_Tr81W", "L_Tr81B",
    "L_Hg01W", "L_Hg01B", "L_Hg11W", "L_Hg11B", "L_Hg21W", "L_Hg21B", "L_Hg31W", "L_Hg31B", "L_Hg41W",
    "L_Hg41B", "L_Hg51W", "L_Hg51B", "L_Hg61W", "L_Hg61B", "L_Hg71W", "L_Hg71B", "L_Hg81W", "L_Hg81B",
    "LDSE01W", "LDSE02W", "LDSE11W", "LDSE12W", "LDSE21W", "LDSE22W", "LDSE31W", "LDSE32W", "LSE01W",
    "LSE02W", "LSE11W", "LSE12W", "LSE21W", "LSE22W", "LSE31W", "LSE32W", "LSE41W", "LSE42W", "LSE51W",
    "LSE52W", "LSE61W", "LSE62W", "LSE71W", "LSE72W", "LSEFin1W", "LSEFin2W", "LGGSW", "LGGSB", "LC0W",
    "LC0B", "LC1W", "LC1B", "LC2W", "LC2B", "LXW", "LXB", "LX0W", "LX0B", "LX1W", "LX1B", "LX2W", "LX2B",
    "LX3W", "LX3B", "LX4W", "LX4B", "LX5W", "LX5B", "LX6W", "LX6B", "LX7W", "LX7B", "LX8W", "LX8B",
    "LX9W", "LX9B", "LX10W", "LX10B", "LX11W", "LX11B", "LX12W", "LX12B"};

static int weight_count(const char* name) {
    // counts implied by the layer shapes of radiancePredict.cu:301-422
    std::string s(name);
    auto ends = [&](const char* suf) { return s.size() >= strlen(suf) && s.compare(s.size() - strlen(suf), strlen(suf), suf) == 0; };
    if (s == "LSEFin1W") return 1200;
    if (s == "LSEFin2W") return 96;
    if (s == "LGGSW") return 24;
    if (s == "LGGSB") return 8;
    if (s == "LC0W") return 16384;
    if (s == "LC0B") return 128;
    if (s == "LC1W") return 8192;
    if (s == "LC1B") return 64;
    if (s == "LC2W") return 2048;
    if (s == "LC2B") return 32;
    if (s == "LXW") return 512;
    if (s == "LXB") return 16;
    if (s == "LX12W") return 16;
    if (s == "LX12B") return 1;
    if (s.rfind("LX", 0) == 0) return ends("W") ? 256 : 16;
    if (s.find("SE") != std::string::npos) return (s[s.size() - 2] == '1') ? 64 : 24;  // SEW0 8x8, SEW1 3x8
    if (s.rfind("LD", 0) == 0) return ends("W") ? 64 : 8;
    // main stack L{,_Tr,_Hg}{0..8}1
    char lvl = s[s.size() - 3];
    int n = (lvl <= '1') ? 8 : (lvl <= '4' ? 16 : 32);
    return ends("W") ? n * n : n;
}

bool net_bind(Net& net, const float* all, int total) {
    int off = 0;
    for (int i = 0; i < 146; ++i) {
        net.p[i] = all + off;
        off += weight_count(kWeightNames[i]);
    }
    return off == total;
}

const float* Net::get(const char* name) const {
    for (int i = 0; i < 146; ++i)
        if (strcmp(kWeightNames[i], name) == 0) return p[i];
    fprintf(stderr, "oracle: unknown weight %s\n", name);
    abort();
}

// ---------------------------------------------------------------------------------------
// MLP operators (radiancePredict.cu:11-238).  T = float restates the device arithmetic;
// T = double is the FP64 evaluation of the same network.
// ---------------------------------------------------------------------------------------
template <class T> static inline T sigm(T x);
// 1/(1+exp(-x)) (radiancePredict.cu:84).  exp of the fp32 network is exp_f32 of fexp.h, the same
// single-precision operations as the device, so host and device agree exactly without double
// arithmetic (the reference's -use_fast_math versions are not reproducible anyway); the FP64
// network uses libm exp.
template <> inline float sigm<float>(float x) { return 1.0f / (1.0f + mrp::exp_f32(-x)); }
template <> inline double sigm<double>(double x) { return 1.0 / (1.0 + exp(-x)); }
template <class T> static inline T expT(T x);
template <> inline float expT<float>(float x) { return mrp::exp_f32(x); }
template <> inline double expT<double>(double x) { return exp(x); }
template <class T> static inline T relu(T x) { return x >= (T)0 ? x : (T)0; }

// Largest |input| of each weight matrix that the MFMA form runs as a matrix product (FP8
// calibration checks): per-thread, switched on by set_amax_recorder.
static thread_local double* t_amax = nullptr;
static thread_local const float* t_amax_w[kAmaxLayers];
void set_amax_recorder(const Net* net, double* amax) {
    t_amax = amax;
    if (!net) return;
    static const char* const tb[3] = {"L", "L_Tr", "L_Hg"}, *db[3] = {"LD", "LD_Tr", "LD_Hg"};
    char nm[32];
    int id = 0;
    for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 9; ++k) {
            snprintf(nm, sizeof nm, "%s%d1W", tb[b], k);
            t_amax_w[id++] = net->get(nm);
        }
    for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 4; ++k) {
            snprintf(nm, sizeof nm, "%s%d1W", db[b], k);
            t_amax_w[id++] = net->get(nm);
        }
    const char* rest[6] = {"LD41W", "LD_Tr41W", "LD_Hg41W", "LSEFin2W", "LC0W", "LC1W"};
    for (const char* r : rest) t_amax_w[id++] = net->get(r);
    t_amax_w[id++] = net->get("LC2W");
    t_amax_w[id++] = net->get("LXW");
    for (int k = 0; k <= 12; ++k) {
        snprintf(nm, sizeof nm, "LX%dW", k);
        t_amax_w[id++] = net->get(nm);
    }
    if (id != kAmaxLayers) abort();
}
template <class T>
static inline void note_input(const float* W, const T* in, int from) {
    if (!t_amax) return;
    for (int id = 0; id < kAmaxLayers; ++id)
        if (t_amax_w[id] == W) {
            for (int j = 0; j < from; ++j) t_amax[id] = std::max(t_amax[id], fabs((double)in[j]));
            return;
        }
}

// AVGMAX (radiancePredict.cu:21-33)
template <class T>
static void avgmax(const T* in, T* out, int n, int read_at, int w1, int w2) {
    out[w1] = (T)0;
    out[w2] = (T)-10000.0;
    for (int i = 0; i < n; ++i) {
        out[w1] += in[read_at + i];
        out[w2] = std::max(out[w2], in[read_at + i]);
    }
    out[w1] /= (T)n;
}
// LinearReLU_SFL / LinearReLU<Size>: bias first, j ascending (radiancePredict.cu:129-224)
template <class T>
static void lin_relu_bf(const T* in, T* out, int from, int to, const float* W, const float* B) {
    note_input(W, in, from);
    for (int i = 0; i < to; ++i) {
        T acc = (T)B[i];
        for (int j = 0; j < from; ++j) acc = mac<T>(acc, in[j], (T)W[j + i * from]);
        out[i] = relu(acc);
    }
}
// untemplated LinearReLU: zero start, bias last (radiancePredict.cu:41-56)
template <class T>
static void lin_relu_bl(const T* in, T* out, int from, int to, const float* W, const float* B) {
    note_input(W, in, from);
    for (int i = 0; i < to; ++i) {
        T acc = (T)0;
        for (int j = 0; j < from; ++j) acc = mac<T>(acc, in[j], (T)W[j + i * from]);
        acc += (T)B[i];
        out[i] = relu(acc);
    }
}
// LinearReLU_Unbias (:58-72) and LinearSm_Unbias (:74-86)
template <class T>
static void lin_relu_nb(const T* in, T* out, int from, int to, const float* W) {
    for (int i = 0; i < to; ++i) {
        T acc = (T)0;
        for (int j = 0; j < from; ++j) acc = mac<T>(acc, in[j], (T)W[j + i * from]);
        out[i] = relu(acc);
    }
}
template <class T>
static void lin_sigm_nb(const T* in, T* out, int from, int to, const float* W) {
    note_input(W, in, from);
    for (int i = 0; i < to; ++i) {
        T acc = (T)0;
        for (int j = 0; j < from; ++j) acc = mac<T>(acc, in[j], (T)W[j + i * from]);
        out[i] = sigm<T>(acc);
    }
}

template <class T>
struct Mlp {
    const Net& net;
    const T *X, *S, *H;
    T A[160], B[160], C[160], A2[160], B2[160], C2[160];
    T pool[75], aw[16], awp[6], tap[8], comb[128];

    Mlp(const Net& n) : net(n) {}

    // SE / SERES macros (radiancePredict.cu:240-279); last < 0 selects SE.
    void block(int last, int from, int to, int n, int pid, const char* sew0, const char* sew1, const char* ww,
               const char* trw, const char* hgw) {
        avgmax<T>(X, pool, n, from, pid + 0, 36 + pid + 0);
        avgmax<T>(S, pool, n, from, pid + 1, 36 + pid + 1);
        avgmax<T>(H, pool, n, from, pid + 2, 36 + pid + 2);
        for (int i = 0; i < 3; ++i) tap[i] = pool[pid + i];
        for (int i = 0; i < 3; ++i) tap[3 + i] = pool[36 + pid + i];
        lin_relu_nb<T>(tap, aw, 8, 8, net.get(sew0));
        lin_sigm_nb<T>(aw, awp, 8, 3, net.get(sew1));
        const T* src[3] = {X, S, H};
        T* in[3] = {A, B, C};
        T* outp[3] = {A2, B2, C2};
        const char* names[3] = {ww, trw, hgw};
        for (int b = 0; b < 3; ++b) {
            for (int i = 0; i < n; ++i) in[b][to + i] = src[b][from + i];
            for (int i = 0; i < n; ++i) in[b][to + i] *= awp[b];
            if (last >= 0)
                for (int i = 0; i < n; ++i) in[b][to + i] = in[b][to + i] + outp[b][last + i];
            char wn[32], bn[32];
            snprintf(wn, sizeof wn, "%sW", names[b]);
            snprintf(bn, sizeof bn, "%sB", names[b]);
            lin_relu_bf<T>(in[b] + to, outp[b] + to, n, n, net.get(wn), net.get(bn));
        }
    }
    void add3(int from, int to, int n) {  // ADD3 (:281-285)
        for (int i = 0; i < n; ++i) {
            A2[to + i] = A2[to + i] + X[from + i];
            B2[to + i] = B2[to + i] + S[from + i];
            C2[to + i] = C2[to + i] + H[from + i];
        }
    }
    void rep3(int from, int to, int n) {  // REP3 (:287-294)
        for (int i = 0; i < n; ++i) { A2[from + i] = A2[to + i]; B2[from + i] = B2[to + i]; C2[from + i] = C2[to + i]; }
        for (int i = 0; i < n; ++i) { A2[to + i] = X[to + i]; B2[to + i] = S[to + i]; C2[to + i] = H[to + i]; }
    }
    // SC(p) tail (:375-409)
    T tail(T srp_p) {
        pool[74] = srp_p;
        for (int i = 0; i < 120; ++i) A[i] = comb[i];
        lin_relu_bl<T>(pool + 72, A + 120, 3, 8, net.get("LGGSW"), net.get("LGGSB"));
        lin_relu_nb<T>(pool, aw, 75, 16, net.get("LSEFin1W"));
        lin_sigm_nb<T>(aw, awp, 16, 6, net.get("LSEFin2W"));
        const int seg_off[6] = {0, 32, 64, 96, 104, 112}, seg_n[6] = {32, 32, 32, 8, 8, 8};
        for (int s = 0; s < 6; ++s)
            for (int i = 0; i < seg_n[s]; ++i) A[seg_off[s] + i] *= awp[s];
        lin_relu_bf<T>(A, B, 128, 128, net.get("LC0W"), net.get("LC0B"));
        lin_relu_bf<T>(B, A, 128, 64, net.get("LC1W"), net.get("LC1B"));
        lin_relu_bf<T>(A, B, 64, 32, net.get("LC2W"), net.get("LC2B"));
        lin_relu_bf<T>(B, A, 32, 16, net.get("LXW"), net.get("LXB"));
        for (int k = 0; k < 6; ++k) {
            char w0[16], b0[16], w1[16], b1[16];
            snprintf(w0, sizeof w0, "LX%dW", 2 * k);
            snprintf(b0, sizeof b0, "LX%dB", 2 * k);
            snprintf(w1, sizeof w1, "LX%dW", 2 * k + 1);
            snprintf(b1, sizeof b1, "LX%dB", 2 * k + 1);
            lin_relu_bf<T>(A, B, 16, 16, net.get(w0), net.get(b0));
            lin_relu_bf<T>(B, C, 16, 16, net.get(w1), net.get(b1));
            for (int i = 0; i < 16; ++i) A[i] = A[i] + C[i];
        }
        lin_relu_bl<T>(A, B, 16, 1, net.get("LX12W"), net.get("LX12B"));
        return B[0];
    }
};

template <class T> static inline T powT(T a, T b);
template <> inline float powT<float>(float a, float b) {  // the only use: scatterrate^4
    return b == 4.0f ? (a * a) * (a * a) : powf(a, b);
}
template <> inline double powT<double>(double a, double b) { return pow(a, b); }

template <class T>
void predict_internal(const Net& net, const T* X_Val, const T* X_Val_Sub, const T* X_Val_Hg, T gamma, T g,
                      const T scatterrate[3], T out[3]) {
    Mlp<T> m(net);
    m.X = X_Val;
    m.S = X_Val_Sub;
    m.H = X_Val_Hg;
    const int DI = 160, P_DI = 24;
    m.tap[6] = g;
    m.tap[7] = gamma;
    T srp[3];
    for (int c = 0; c < 3; ++c) srp[c] = powT<T>(scatterrate[c], (T)4.0);
    m.pool[72] = g;
    m.pool[73] = gamma;
    // radiancePredict.cu:332-373
    m.block(-1, 0, 0, 8, 0, "LSE01W", "LSE02W", "L01", "L_Tr01", "L_Hg01");
    m.add3(0, 0, 8);
    m.block(0, 8, 8, 8, 3, "LSE11W", "LSE12W", "L11", "L_Tr11", "L_Hg11");
    m.rep3(0, 8, 8);
    m.block(0, 16, 16, 16, 6, "LSE21W", "LSE22W", "L21", "L_Tr21", "L_Hg21");
    m.add3(16, 16, 16);
    m.block(16, 32, 32, 16, 9, "LSE31W", "LSE32W", "L31", "L_Tr31", "L_Hg31");
    m.add3(32, 32, 16);
    m.block(32, 48, 48, 16, 12, "LSE41W", "LSE42W", "L41", "L_Tr41", "L_Hg41");
    m.rep3(32, 48, 16);
    m.block(32, 64, 64, 32, 15, "LSE51W", "LSE52W", "L51", "L_Tr51", "L_Hg51");
    m.add3(64, 64, 32);
    m.block(64, 96, 96, 32, 18, "LSE61W", "LSE62W", "L61", "L_Tr61", "L_Hg61");
    m.add3(96, 96, 32);
    m.block(96, 128, 128, 32, 21, "LSE71W", "LSE72W", "L71", "L_Tr71", "L_Hg71");
    m.add3(128, 128, 32);
    lin_relu_bf<T>(m.A2 + 128, m.comb + 0, 32, 32, net.get("L81W"), net.get("L81B"));
    lin_relu_bf<T>(m.B2 + 128, m.comb + 32, 32, 32, net.get("L_Tr81W"), net.get("L_Tr81B"));
    lin_relu_bf<T>(m.C2 + 128, m.comb + 64, 32, 32, net.get("L_Hg81W"), net.get("L_Hg81B"));
    m.block(-1, DI + 0, 0, 8, P_DI + 0, "LDSE01W", "LDSE02W", "LD01", "LD_Tr01", "LD_Hg01");
    m.add3(DI + 0, 0, 8);
    m.block(0, DI + 8, 8, 8, P_DI + 3, "LDSE11W", "LDSE12W", "LD11", "LD_Tr11", "LD_Hg11");
    m.add3(DI + 8, 8, 8);
    m.block(8, DI + 16, 16, 8, P_DI + 6, "LDSE21W", "LDSE22W", "LD21", "LD_Tr21", "LD_Hg21");
    m.add3(DI + 16, 16, 8);
    m.block(16, DI + 24, 24, 8, P_DI + 9, "LDSE31W", "LDSE32W", "LD31", "LD_Tr31", "LD_Hg31");
    m.add3(DI + 24, 24, 8);
    lin_relu_bf<T>(m.A2 + 24, m.comb + 96, 8, 8, net.get("LD41W"), net.get("LD41B"));
    lin_relu_bf<T>(m.B2 + 24, m.comb + 104, 8, 8, net.get("LD_Tr41W"), net.get("LD_Tr41B"));
    lin_relu_bf<T>(m.C2 + 24, m.comb + 112, 8, 8, net.get("LD_Hg41W"), net.get("LD_Hg41B"));
    // radiancePredict.cu:411-421
    T P[3];
    P[0] = m.tail(srp[0]);
    P[1] = (scatterrate[1] == scatterrate[0]) ? P[0] : m.tail(srp[1]);
    if (scatterrate[2] == scatterrate[0]) P[2] = P[0];
    else if (scatterrate[2] == scatterrate[1]) P[2] = P[1];
    else P[2] = m.tail(srp[2]);
    for (int c = 0; c < 3; ++c) {
        T e = expT<T>(P[c]) - (T)1.0;
        out[c] = std::max(e, (T)0) * srp[c];
    }
}
template void predict_internal<float>(const Net&, const float*, const float*, const float*, float, float,
                                      const float*, float*);
template void predict_internal<double>(const Net&, const double*, const double*, const double*, double, double,
                                       const double*, double*);

// ---------------------------------------------------------------------------------------
// Camera ray (volume.cu:198-216; camera.cu:166-192) and NNPredict (render.cuh:808-875)
// ---------------------------------------------------------------------------------------
PixelState pixel_state(const Volume& v, const View& view, int row, int col) {
    PixelState st;
    uint32_t pix = (uint32_t)(row * view.width + col);
    float xu, xv;
    if (view.jitter < 0) {
        mrp::PhiloxStream rs(view.seed, view.frame, pix, mrp::RNG_JITTER, view.view_index);
        xu = rs.next();
        xv = rs.next();
    } else {
        xu = xv = view.jitter;
    }
    float u = 1.0f - ((float)col + xu) / (float)view.width;
    float vv = 1.0f - ((float)row + xv) / (float)view.height;
    float sy = view.tan_half_fov;
    float sx = view.tan_half_fov * ((float)view.width / (float)view.height);
    f3 fwd = normalize(-view.ori);
    f3 dir = fwd + (view.right * ((u * 2.0f - 1.0f) * sx)) + (view.up * ((vv * 2.0f - 1.0f) * sy));
    dir = normalize(dir);          // volume.cu:216
    dir = normalize(dir);          // render.cuh:811
    f3 light = normalize(view.light_dir);  // render.cuh:812
    MarchResult m = get_sample_point(v, view.ori, dir, view.alpha);
    Basis b = basis_from_normal(dir, view.seed, view.frame, pix, view.view_index);
    st.active = m.active;
    st.pos = m.pos;
    st.steps = m.steps;
    st.XMain = b.x;
    st.YMain = b.y;
    st.ZMain = b.z;
    st.LXMain = normalize(light);  // GetMatrixFromNormal normalises again (render.cuh:696)
    st.LightDir = light;
    st.dir = dir;
    return st;
}

void predict_pixel(const Volume& v, const Net& net, const View& view, int row, int col, float out4[4],
                   float* feats_out, int* steps_out) {
    PixelState st = pixel_state(v, view, row, col);
    if (steps_out) *steps_out = st.steps;
    float feats[NFEAT];
    gather_features(v, st.active, st.pos, st.XMain, st.YMain, st.ZMain, st.LXMain, st.LightDir, view.g, view.alpha,
                    feats);
    if (feats_out) memcpy(feats_out, feats, sizeof feats);
    if (!st.active) {  // SkyBox(dir), distance -1 -> max(0,.) = 0 (render.cuh:829,874; volume.cu:224)
        const f3 sky = sky_box(v, st.dir);
        out4[0] = fmaxf(0.0f, sky.x);
        out4[1] = fmaxf(0.0f, sky.y);
        out4[2] = fmaxf(0.0f, sky.z);
        out4[3] = 0.0f;
        return;
    }
    f3 sr = view.scatter_rate / 1.001f;  // render.cuh:842
    float srv[3] = {sr.x, sr.y, sr.z}, rgb[3];
    predict_internal<float>(net, feats, feats + 2 * PC, feats + PC, feats[3 * PC], view.g, srv, rgb);
    f3 res = (mk(rgb[0], rgb[1], rgb[2]) * view.light_color) * view.scatter_rate;  // render.cuh:874
    float w = length(view.ori - st.pos);
    out4[0] = fmaxf(0.0f, res.x);
    out4[1] = fmaxf(0.0f, res.y);
    out4[2] = fmaxf(0.0f, res.z);
    out4[3] = fmaxf(0.0f, w);
}

// ---------------------------------------------------------------------------------------
// Mip pyramid (volume.cu:1474-1515 with ParallelFill omp.cpp:4-18 and Sample :845-863)
// ---------------------------------------------------------------------------------------
static void resample(const float* src, int sres, float* dst, int res) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < res; ++i)
        for (int j = 0; j < res; ++j)
            for (int k = 0; k < res; ++k) {
                float u = (float)((i + 0.5) / res), vv = (float)((j + 0.5) / res), w = (float)((k + 0.5) / res);
                dst[((size_t)i * res + j) * res + k] = sample_border(src, sres, mk(u, vv, w), true);
            }
}
void build_pyramid(const float* grid, int N, float* const mips[9]) {
    int res = 256;
    while (res * 2 < N) res <<= 1;
    const float* source = grid;
    int sres = N;
    std::vector<float> tmp_a, tmp_b;
    bool use_a = true;
    for (; res > 256; res >>= 1) {
        std::vector<float>& t = use_a ? tmp_a : tmp_b;
        t.assign((size_t)res * res * res, 0.0f);
        resample(source, sres, t.data(), res);
        source = t.data();
        sres = res;
        use_a = !use_a;
    }
    for (int m = 0; m < 9; ++m) {
        int r = 256 >> m;
        resample(source, sres, mips[m], r);
        source = mips[m];
        sres = r;
    }
}

// ---------------------------------------------------------------------------------------
// Transmittance mips Fill_TR (volume.cu:1535-1560) launched per level (volume.cu:1636-1653)
// ---------------------------------------------------------------------------------------
void bake_tr(const float* const mips[9], f3 light_dir, float alpha, float* const tr[8]) {
    for (int m = 0; m < 8; ++m) {
        int res = 128 >> m;
        float a = alpha / powf(1.73f, (float)m + 1.0f);
        const float* src = mips[m + 1];
        int sres = 256 >> (m + 1);
#pragma omp parallel for schedule(static)
        for (int x = 0; x < res; ++x)
            for (int y = 0; y < res; ++y)
                for (int z = 0; z < res; ++z) {
                    f3 uv = (mk((float)x, (float)y, (float)z) + 0.5f) / (float)res;
                    f3 ori = uv - 0.5f;
                    float dis = ray_box_distance(ori, light_dir);
                    float st = dis / 128.0f;
                    f3 lp = ori;
                    float sum = 0.0f;
                    for (int i = 0; i < 128; ++i) {
                        lp = lp + light_dir * st;
                        sum = sum + sample_border(src, sres, lp + 0.5f, false);
                    }
                    tr[m][((size_t)x * res + y) * res + z] = expf(-sum * a * st) * 1.0f;
                }
    }
}

// ---------------------------------------------------------------------------------------
// HG LUT (volume.cu:1186-1400): scrambled Halton + per-texel jitter (Philox here)
// ---------------------------------------------------------------------------------------
float halton_scrambled(unsigned dim, unsigned a) {
    const unsigned base = mrp::kHaltonPrimes[dim], offset = mrp::kHaltonPrimeSums[dim];
    const float inv_base = 1.0f / (float)base;
    unsigned rev = 0;
    float inv_base_n = 1.0f;
    while (a) {
        unsigned next = a / base;
        unsigned digit = a - next * base;
        rev = rev * base + mrp::kHaltonPerm[offset + digit];
        inv_base_n *= inv_base;
        a = next;
    }
    float r = inv_base_n * ((float)rev + inv_base * (float)mrp::kHaltonPerm[offset] / (1.0f - inv_base));
    return std::min(r, 0.99999994f);
}
static inline float frac_(float f) { return f - (float)(long)f; }  // vector.cu:20

void hg_lut_qmc(float g, uint32_t seed, float* lut) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < LUT; ++j)
        for (int i = 0; i < LUT; ++i) {
            // volume.cu:1330-1332 (double there); the library's single-precision form
            const float costheta = ((float)i + 0.5f) / (float)LUT * 2.0f - 1.0f;
            const float aa = ((float)j + 0.5f) / (float)LUT;
            const float angle = aa * 1.04719755f;
            mrp::PhiloxStream rs(seed, 0u, (uint32_t)(i + j * LUT), mrp::RNG_LUT, 0u);
            float bn = rs.next();
            float hg = 0.0f;
            float R = tanf(angle);
            f3 view = mk(1, 0, 0);
            f3 D = mk(costheta, sqrtf(1.0f - costheta * costheta), 0.0f);
            float mis = 0;
            {
                float p = R * R - 1.0f + D.x * D.x;
                if (p > 0) {
                    float sdotl = g > 0 ? D.x : -D.x;
                    float w = sqrtf(p);
                    float d0 = fmaxf(0.0f, sdotl - w) / R;
                    float d1 = (sdotl + w) / R;
                    float d = d1 * d1 * d1 - d0 * d0 * d0;
                    mis = (float)(d1 > d0) * d;
                }
                mis = fminf(fmaxf(mis, 1e-8f), 1.0f);
                float a = fabsf(g);
                a = (double)a < 0.6 ? 0.0f
                    : ((double)a < 0.7 ? (float)(((double)a - 0.6) * 9)
                                       : (float)(pow(((double)a - 0.7) / 0.3, 0.05) * 0.1 + 0.9));
                mis = 1.0f - 1.0f / (1.0f + a * mis / (1.0f - a));
            }
            for (int k = 0; k < 1024; ++k) {
                if (frac_(halton_scrambled(4, k) + bn) < mis) {
                    float rx = frac_(halton_scrambled(0, k) + bn), ry = frac_(halton_scrambled(1, k) + bn);
                    f3 l = sample_hg(rx, ry, view, g);
                    float sdotl = dot(D, l);
                    float p = R * R - 1.0f + sdotl * sdotl;
                    if (p <= 0) continue;
                    float w = sqrtf(p);
                    float d0 = sdotl - w;
                    float d1 = (sdotl + w) / R;
                    d0 = fmaxf(0.0f, d0) / R;
                    float d = d1 * d1 * d1 - d0 * d0 * d0;
                    hg += fmaxf(0.0f, d / 4.0f);
                } else {
                    float rx = frac_(halton_scrambled(0, k) + bn), ry = frac_(halton_scrambled(1, k) + bn),
                          rz = frac_(halton_scrambled(2, k) + bn);
                    const float theta = rx * 6.28318531f;
                    const float phi = acosf(2.0f * ry - 1.0f);
                    float r = powf(rz, 1.0f / 3.0f);
                    float st = sinf(theta), ct = cosf(theta), sp = sinf(phi), cp = cosf(phi);
                    float x = R * r * sp * ct, y = R * r * sp * st, z = R * r * cp;
                    f3 rp = mk(x, y, z) + D;
                    float c = dot(normalize(rp), view);
                    hg += henyey_greenstein(c, g);
                }
            }
            hg /= 1024.0f;
            lut[j * LUT + i] = hg;
        }
}

// FP64 reference value of the LUT texel: mean of HG_ref over a ball of radius tan(angle)
// centred at unit distance in direction (costheta, sin, 0), seen along (1,0,0).  Integrates
// over directions l inside the ball's cone: mean = (1/4pi) Int HG(l.v) (d1^3 - d0^3) dOmega,
// d0/d1 = chord entry/exit distances along l divided by the radius (the chord-weight form
// the reference's importance branch uses, volume.cu:1369-1377).  Gauss-Legendre in both
// angles around the ball centre direction.
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1, p2 = 0;
            for (int j = 0; j < n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-15) break;
        }
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * pp * pp);
    }
}
double hg_ball_mean_fp64(double costheta, double angle, double g, int n_psi, int n_phi) {
    double R = tan(angle);
    double st = sqrt(std::max(0.0, 1 - costheta * costheta));
    double c[3] = {costheta, st, 0};             // ball centre direction
    double e1[3] = {-st, costheta, 0}, e2[3] = {0, 0, 1};  // orthonormal frame around c
    double psi_max = R < 1 ? asin(R) : M_PI;
    std::vector<double> xs, ws, xp, wp;
    gauss_legendre(n_psi, xs, ws);
    gauss_legendre(n_phi, xp, wp);
    double sum = 0;
    for (int a = 0; a < n_psi; ++a) {
        double psi = 0.5 * psi_max * (xs[a] + 1), wpsi = 0.5 * psi_max * ws[a];
        double sc = cos(psi), ss = sin(psi);
        // chord of the ball along l: t^2 - 2 t cos(psi) + 1 - R^2 = 0
        double p = R * R - 1 + sc * sc;
        if (p <= 0) continue;
        double w = sqrt(p);
        double d0 = std::max(0.0, sc - w) / R, d1 = (sc + w) / R;
        double chord = d1 * d1 * d1 - d0 * d0 * d0;
        double inner = 0;
        for (int b = 0; b < n_phi; ++b) {
            double phi = M_PI * (xp[b] + 1), wphi = M_PI * wp[b];
            double l0 = sc * c[0] + ss * (cos(phi) * e1[0] + sin(phi) * e2[0]);
            double g2 = g * g;
            double hg = 0.25 * (1 - g2) / pow(1 + g2 - 2 * g * l0, 1.5);
            inner += wphi * hg;
        }
        sum += wpsi * ss * chord * inner;
    }
    return sum / (4 * M_PI);
}

// ---------------------------------------------------------------------------------------
// Path tracer: Tr (render.cuh:60-103), DeterminateNextVertex (:399-434),
// CalculateRadiance (:535-597); FP64 Welford per pixel.
// ---------------------------------------------------------------------------------------
float tr_stochastic(const Volume& v, mrp::PhiloxStream& rs, f3 pos, f3 dir, float dis, float alpha, float tr_scale) {
    float smax = v.max_density * alpha;
    float tr = 1.0f, t = 0.0f;
    int loop = 0;
    while (loop++ < 10000 && tr > 0.5f) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax * tr_scale;
        if (t > dis) break;
        float d = density(v, pos + (dir * t));
        float S = d * alpha;
        tr *= 1.0f - fmaxf(0.0f, S / smax);
        if (d < 0) t -= d;
    }
    while (loop++ < 10000 && (double)tr > 0.001) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax * tr_scale;
        if (t > dis) break;
        float d = density(v, pos + (dir * t));
        float S = d * alpha;
        if (rs.next() < fmaxf(0.0f, S / smax)) { tr = 0.0f; break; }
        if (d < 0) t -= d;
    }
    return tr;
}

static bool next_vertex(const Volume& v, mrp::PhiloxStream& rs, float alpha, float g, f3 pos, f3 dir, float dis,
                        f3* np, f3* nd, long long* taps) {
    float smax = v.max_density * alpha;
    float t = 0;
    int loop = 0;
    while (loop++ < 10000) {
        float rk = rs.next();
        t -= logf(1.0f - rk) / smax;
        if (t > dis) return false;
        rk = rs.next();
        float d = density(v, pos + (dir * t));
        ++*taps;
        float S = d * alpha;
        if (S / smax > rk) break;
        if (d < 0) t -= d;
    }
    float e0 = rs.next();
    float e1 = rs.next();
    *nd = sample_hg(e0, e1, dir, g);
    *np = (dir * t) + pos;
    return true;
}

void pt_pixel(const Volume& v, const View& view, const PTParams& pp, int row, int col, double* mean3, double* m2_3,
              long long* taps) {
    uint32_t pix = (uint32_t)(row * view.width + col);
    double mean[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
    f3 light = normalize(view.light_dir);
    for (int s = 0; s < pp.spp; ++s) {
        mrp::PhiloxStream rs(view.seed, view.frame, pix, mrp::RNG_PT, (uint32_t)s);
        float xu = view.jitter, xv = view.jitter;
        if (view.jitter < 0) { xu = rs.next(); xv = rs.next(); }
        float u = 1.0f - ((float)col + xu) / (float)view.width;
        float vv = 1.0f - ((float)row + xv) / (float)view.height;
        float sy = view.tan_half_fov, sx = view.tan_half_fov * ((float)view.width / (float)view.height);
        f3 fwd = normalize(-view.ori);
        f3 dir = fwd + (view.right * ((u * 2.0f - 1.0f) * sx)) + (view.up * ((vv * 2.0f - 1.0f) * sy));
        dir = normalize(dir);
        dir = normalize(dir);
        f3 res = mk(0, 0, 0);
        float dis = ray_box_offset(view.ori, dir);
        if (dis < 0) res = sky_box(v, dir);  // render.cuh:545-546
        if (dis >= 0) {
            f3 spos = view.ori + dir * dis, rdir = dir, csr = mk(1, 1, 1);
            for (int b = 0; b < pp.max_bounce; ++b) {
                f3 npos, ndir;
                float maxd = ray_box_distance(spos, rdir);
                bool in = next_vertex(v, rs, view.alpha, view.g, spos, rdir, maxd, &npos, &ndir, taps);
                if (!in) {  // render.cuh:569-573
                    res = res + sky_box(v, rdir) * csr;
                    break;
                }
                spos = npos;
                csr = csr * view.scatter_rate;
                float ld = ray_box_distance(spos, light);
                float phase = henyey_greenstein(dot(rdir, light), view.g);
                float tr = pp.nee_baked == 1   ? mip_tr(v, 0, spos)
                           : pp.nee_baked == 2 ? sample_clamp(pp.ctr, pp.ctr_res, spos + 0.5f)
                                               : tr_stochastic(v, rs, spos, light, ld, view.alpha, pp.tr_scale);
                res = res + view.light_color * csr * (tr * phase);
                rdir = ndir;
            }
        }
        double x[3] = {res.x, res.y, res.z};
        for (int c = 0; c < 3; ++c) {
            double d = x[c] - mean[c];
            mean[c] += d / (double)(s + 1);
            m2[c] += d * (x[c] - mean[c]);
        }
    }
    for (int c = 0; c < 3; ++c) { mean3[c] = mean[c]; m2_3[c] = m2[c]; }
}

float cached_tr_texel(const Volume& v, int res, long long i, f3 light, float alpha, float tr_scale, uint32_t seed) {
    const int z = (int)(i % res), y = (int)((i / res) % res), x = (int)(i / ((long long)res * res));
    const f3 pos = (mk((float)x, (float)y, (float)z) + 0.5f) / (float)res - 0.5f;  // volume.cu:1570-1573
    const int k = density(v, pos) > 0.0f ? 5 : 2;                                   // :1574
    const float dis = ray_box_distance(pos, light);
    float acc = 0.0f;
    for (int e = 0; e < k; ++e) {
        mrp::PhiloxStream rs(seed, 1u, (uint32_t)i, mrp::RNG_BAKE, (uint32_t)e);
        acc += tr_stochastic(v, rs, pos, light, dis, alpha, tr_scale);
    }
    return acc / (float)k;  // TR_MUL = 1 (volume.cu:885)
}

f3 point_sample(const Volume& v, mrp::PhiloxStream& rs, f3 ori, f3 dir, f3 light, f3 light_color, float scatter,
                float alpha, float g, int max_bounce, float tr_scale) {
    f3 res = mk(0, 0, 0);
    const float e0 = rs.next();
    const float e1 = rs.next();
    f3 spos = ori, rdir = sample_hg(e0, e1, dir, g), csr = mk(1, 1, 1);  // render.cuh:623-624
    long long taps = 0;
    for (int b = 0; b < max_bounce; ++b) {
        f3 npos, ndir;
        const float dis = ray_box_distance(spos, rdir);
        if (!next_vertex(v, rs, alpha, g, spos, rdir, dis, &npos, &ndir, &taps)) {  // :631-636
            res = res + sky_box(v, rdir) * csr;
            break;
        }
        spos = npos;
        csr = (csr * 1.001f) * scatter;  // :640
        const float ld = ray_box_distance(spos, light);
        const float phase = henyey_greenstein(dot(rdir, light), g);
        res = res + light_color * csr * (tr_stochastic(v, rs, spos, light, ld, alpha, tr_scale) * phase);  // :642-644
        rdir = ndir;
    }
    return res;
}

// ---------------------------------------------------------------------------------------
// Tile feature gather
// ---------------------------------------------------------------------------------------
// One root of the trapezoid optical depth inside a step (render.cuh:223-241): A u^2 + B u + C = 0
static float trapezoid_offset(float A, float B, float C, float dt, float S, float lastS, float delta_tau) {
    if (fabsf(A) < 1e-6f) return delta_tau / (dt * (S + lastS) * 0.5f);  // linear fallback
    float D = B * B - 4.0f * A * C;
    D = fmaxf(D, 0.0f);
    float sqrtD = sqrtf(D);
    float q = -0.5f * (B + sqrtD);  // the reference's double -0.5 scales exactly
    float u1 = q / A;
    float u2 = C / q;
    return u1 == fminf(fmaxf(u1, 0.0f), 1.0f) ? u1 : u2;  // saturate
}
// CheckExitCondition (render.cuh:199-288)
static bool check_exit(float S, float lastS, float dt, float t, float& tau, float& tau_target, Ts4& ts,
                       bool& first_hit) {
    float tau_next = tau + (S + lastS) * 0.5f * dt;
    if (tau_next >= tau_target) {
        float delta_tau = tau_target - tau;
        float A = 0.5f * dt * (S - lastS);
        float B = dt * lastS;
        float C = -delta_tau;
        float offset = trapezoid_offset(A, B, C, dt, S, lastS, delta_tau);
        offset = fminf(fmaxf(offset, 0.0f), 1.0f);
        float hit_t = t - dt * (1.0f - offset);
        if (first_hit) {
            ts.x = hit_t;
            tau_target = kTauSecond;
            first_hit = false;
            if (tau_next >= tau_target) {
                float delta_tau2 = tau_target - tau;
                float offset2 = trapezoid_offset(A, B, -delta_tau2, dt, S, lastS, delta_tau2);  // not clamped
                ts.y = t - dt * (1.0f - offset2);
                return true;
            }
        } else {
            ts.y = hit_t;
            return true;
        }
    }
    tau = tau_next;
    return false;
}
// DeterminateEstimatedHit, float4 version (render.cuh:290-397)
Ts4 estimated_hit_tau(const Volume& v, float alpha, f3 pos, f3 dir, float dis) {
    float smax = v.max_density * alpha;
    float t = 0.0f;
    Ts4 ts = {0.0f, 0.0f, FLT_MAX, 0.0f};
    float last_non_empty_t = 0.0f;
    float t_step = 1.0f / (float)v.N;
    if (smax * t_step > kTauStepMax) t_step = kTauStepMax / smax;
    float tau = 0.0f, tau_target = kTauFirst;
    bool first_hit = true;
    int loop = 0;
    float last_dens = 0.0f;
    while (loop++ < 10000) {
        if (t > dis) {
            if (first_hit) ts.x = last_non_empty_t;
            else ts.y = last_non_empty_t;
            break;
        }
        float d = density(v, pos + (dir * t));
        if (d > 0.0f) {
            float S = d * alpha;
            if (check_exit(S, last_dens, t_step, t, tau, tau_target, ts, first_hit)) break;
            ts.z = fminf(t, ts.z);
            t += t_step;
            last_dens = S;
        } else {
            if (last_dens > 0.0f) {
                float st = t_step + d;  // the negative SDF value shortens the step
                if (st > 0.0f) {
                    if (check_exit(0.0f, last_dens, st, t, tau, tau_target, ts, first_hit)) break;
                    last_non_empty_t = t + st;
                } else {
                    last_non_empty_t = t;
                }
            }
            t += fmaxf(-d * 1.0f, 0.5f / (float)v.N);
            last_dens = 0.0f;
        }
    }
    if (ts.z == FLT_MAX) ts.z = 0.0f;
    float survival = expf(-tau);
    ts.w = fminf(fmaxf(1.0f - survival, 0.0f) / 0.75f, 1.0f);
    return ts;
}
// GetSampleT (render.cuh:464-486)
Ts4 get_sample_t(const Volume& v, f3 ori, f3 dir, float alpha) {
    float dis = ray_box_offset(ori, dir);
    if (dis < 0.0f) return Ts4{0.0f, 0.0f, 0.0f, 0.0f};
    f3 sp = ori + dir * dis;
    float max_dis = ray_box_distance(sp, dir);
    Ts4 ts = estimated_hit_tau(v, alpha, sp, dir, max_dis);
    if (ts.x > 0.0f) {
        ts.x += dis;
        ts.z += dis;
    }
    if (ts.y > 0.0f) ts.y += dis;
    return ts;
}

bool tile_gather_pixel(const Volume& v, const View& view, int row, int col, int off_x, int off_y, float* out) {
    const int W = view.width, H = view.height;
    // GatherSamples (volume.cu:257-283): sub-pixel offset in eighths, centre of the pixel
    float u = 1.0f - ((float)col + (float)off_x / 8.0f + 0.5f) / (float)W;
    float vv = 1.0f - ((float)row + (float)off_y / 8.0f + 0.5f) / (float)H;
    float sy = view.tan_half_fov;
    float sx = view.tan_half_fov * ((float)W / (float)H);
    f3 fwd = normalize(-view.ori);
    f3 dir = fwd + (view.right * ((u * 2.0f - 1.0f) * sx)) + (view.up * ((vv * 2.0f - 1.0f) * sy));
    f3 right_off = (-(view.right * sx)) / ((float)W * 4.0f);
    f3 up_off = (-(view.up * sy)) / ((float)H * 4.0f);
    f3 L = normalize(view.light_dir);
    float* tgrid = out + 3 * PC + 5;
    for (int i = 0; i < TILE_ROW; ++i) out[i] = 0.0f;
    // NNGatherSamples (render.cuh:1049-1103): 16 groups of 2x2 sub-rays
    struct Sample {
        f3 d;
        float w;
    } samples[16];
    f3 avg = mk(0, 0, 0);
    float minx = FLT_MAX, miny = FLT_MAX, maxx = -FLT_MAX, maxy = -FLT_MAX;
    float hor = 0.0f, vert = 0.0f;
    auto prepare = [&](float ox, float oy, bool do25, int gid, float& h, float& vt) {  // PrepareRay
        f3 sd = dir + right_off * ox + up_off * oy;
        sd = normalize(sd);
        Ts4 ts = get_sample_t(v, view.ori, sd, view.alpha);
        Sample r{sd, do25 ? ts.y : ts.x};
        const float rough = 1.0f / 3.0f;
        float min_t = ts.z * rough;
        float max_t = fmaxf(ts.x, ts.y) * rough;
        tgrid[gid] = (min_t + max_t) * 0.5f;
        tgrid[gid + 1] = max_t - min_t;
        tgrid[gid / 2 + 128] = ts.w;
        if (r.w != 0.0f) {
            minx = fminf(minx, ox);
            miny = fminf(miny, oy);
            maxx = fmaxf(maxx, ox);
            maxy = fmaxf(maxy, oy);
            avg = avg + ((ts.y > 0.0f) ? sd * 2.0f : sd);
            float sum_t = ts.x + ts.y;
            h += sum_t;
            vt += sum_t;
        }
        return r;
    };
    int samp = 0, gid = 0, cb = 0;
    for (int y = -3; y <= 3; y += 2, cb = (cb + 1) & 1, gid += 16) {
        for (int x = -3; x <= 3; x += 2, ++samp, gid += 4) {
            bool do25 = ((samp + cb) & 1) != 0;
            float h = 0.0f, vt = 0.0f;
            Sample t = prepare((float)x - 0.5f, (float)y - 0.5f, do25, gid, h, vt);
            const float ox[3] = {(float)x + 0.5f, (float)x - 0.5f, (float)x + 0.5f};
            const float oy[3] = {(float)y - 0.5f, (float)y + 0.5f, (float)y + 0.5f};
            const int go[3] = {2, 16, 18};
            for (int k = 0; k < 3; ++k) {
                Sample c = prepare(ox[k], oy[k], do25, gid + go[k], h, vt);
                if ((c.w != 0.0f && ((c.w < t.w) != do25)) || t.w == 0.0f) t = c;
            }
            hor += (x > 0) ? h : -h;
            vert += (y > 0) ? vt : -vt;
            int fill = samp / 2 + (do25 ? 8 : 0);
            samples[fill] = t;
        }
    }
    if (minx == FLT_MAX) {
        out[0] = -1000.0f;
        return false;
    }
    // GetMatrixFromNormal(seed, avgDir, 0) (render.cuh:694-708): with zero spread the cap
    // direction is UniformSampleSphere(0, 0) = (0, 0, 1); the reference loops forever when
    // |dot| < 1e-5, here the basis then keeps (0, 1, 0), (0, 0, 1)
    f3 X = normalize(avg), Y = mk(0, 1, 0), Z = mk(0, 0, 1);
    {
        f3 r = uniform_sample_sphere(0.0f, 0.0f);
        if (!((double)fabsf(dot(r, X)) < 0.00001)) {
            Y = normalize(cross(X, r));
            Z = cross(Y, X);
        }
    }
    // gamma outputs (render.cuh:1114-1134)
    {
        minx -= 0.5f;
        miny -= 0.5f;
        maxx += 0.5f;
        maxy += 0.5f;
        f3 bl = normalize(dir + right_off * minx + up_off * miny);
        f3 rc = normalize(dir + right_off * maxx);
        f3 uc = normalize(dir + up_off * maxy);
        auto sc = [](float x) { return fmaxf(-1.0f, fminf(1.0f, x)); };
        float* g = out + 3 * PC;
        g[0] = acosf(sc(dot(X, L)));
        float gbl = acosf(sc(dot(bl, L)));
        g[1] = acosf(sc(dot(rc, L))) - gbl;
        g[2] = acosf(sc(dot(uc, L))) - gbl;
        g[3] = hor;
        g[4] = vert;
    }
    // SetupInputData2 (radiancePredict.cu:487-547)
    float* Xv = out;
    float* Xh = out + PC;
    float* Xs = out + 2 * PC;
    for (int i = 0; i < 16; ++i) {
        const Sample& sm = samples[i];
        if (sm.w != 0.0f) {
            f3 sp = view.ori + sm.d * sm.w;
            Xv[i] = logf(1.0f + view.alpha * mip_density(v, 0, sp) / 64.0f);
            Xh[i] = logf(henyey_greenstein(dot(sm.d, L), view.g) + 1.0f);
            Xs[i] = shadow_term_tr(v, sp, L, 0);
        } else {
            Xv[i] = Xh[i] = Xs[i] = 0.0f;
        }
    }
    int sid = 0;
    for (int i = 16; i < PC; ++i) {
        while (samples[sid].w == 0.0f) sid = (sid + 1) & 15;
        const Sample& sm = samples[sid];
        sid = (sid + 1) & 15;
        f3 sp = view.ori + sm.d * sm.w;
        StencilPoint s = get_samples23(i);
        float layer = (float)s.mip + 0.1f;
        int mip = (int)layer;
        f3 p;
        if (s.type >= 4) {
            f3 Ld = stencil_dir(i);
            p = sp + ((X * Ld.x + Y * Ld.y + Z * Ld.z) * s.offset);
        } else {
            p = sp + L * s.offset;
        }
        Xv[i] = logf(1.0f + view.alpha * mip_density(v, mip, p) / 64.0f);
        f3 ms = normalize(p - sp);
        float radius = (float)(1 << mip) / 256.0f;
        float angle = atanf(0.5f * radius / s.offset);
        float c1 = dot(X, ms), c2 = dot(L, ms);
        float uu = c1 * 0.5f + 0.5f, u2 = c2 * 0.5f + 0.5f;
        float lv = angle / (3.1415926535f * 60.0f / 180.0f);
        float h0 = sample_lut(v.lut, uu, lv), h1 = sample_lut(v.lut, u2, lv);
        Xh[i] = logf(h0 * h1 + 1.0f);
        Xs[i] = shadow_term_tr(v, p, L, mip - 1 > 0 ? mip - 1 : 0);
    }
    return true;
}

}  // namespace oracle
