// extern "C" entry points of the CPU oracle, for ctypes (tests/, smoke(), bench.py's
// cpu_baseline leg).  TEST INFRASTRUCTURE ONLY.
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../mrpnn_amd/csrc/fbm.h"
#include "mrp_oracle.hpp"

using namespace oracle;

// Same layout as mrp_view in include/mrp_c_api.h.
struct orc_view {
    float ori[3], up[3], right[3], light_dir[3], light_color[3], scatter_rate[3];
    float alpha, g, tan_half_fov, jitter;
    int32_t width, height;
    uint32_t seed, frame, view_id;
    int32_t volume_index;
};

static std::vector<float> g_weights;
static Net g_net;
static bool g_net_ok = false;

static View to_view(const orc_view* v) {
    View w;
    w.ori = mk(v->ori[0], v->ori[1], v->ori[2]);
    w.up = mk(v->up[0], v->up[1], v->up[2]);
    w.right = mk(v->right[0], v->right[1], v->right[2]);
    w.light_dir = mk(v->light_dir[0], v->light_dir[1], v->light_dir[2]);
    w.light_color = mk(v->light_color[0], v->light_color[1], v->light_color[2]);
    w.scatter_rate = mk(v->scatter_rate[0], v->scatter_rate[1], v->scatter_rate[2]);
    w.alpha = v->alpha;
    w.g = v->g;
    w.tan_half_fov = v->tan_half_fov;
    w.width = v->width;
    w.height = v->height;
    w.seed = v->seed;
    w.frame = v->frame;
    w.view_index = v->view_id;
    w.jitter = v->jitter;
    return w;
}

static size_t mip_offset(int m) {
    size_t o = 0;
    for (int i = 0; i < m; ++i) o += (size_t)(256 >> i) * (256 >> i) * (256 >> i);
    return o;
}
static size_t tr_offset(int m) {
    size_t o = 0;
    for (int i = 0; i < m; ++i) o += (size_t)(128 >> i) * (128 >> i) * (128 >> i);
    return o;
}
static std::vector<float> g_env;
static int g_env_w = 0, g_env_h = 0;
static float g_env_exp = 1.0f;

static Volume to_volume(const float* grid, int N, float max_density, const float* mips, const float* tr,
                        const float* lut) {
    Volume v;
    v.env = g_env.empty() ? nullptr : g_env.data();
    v.env_w = g_env_w;
    v.env_h = g_env_h;
    v.env_exp = g_env_exp;
    v.N = N;
    v.grid = grid;
    v.max_density = max_density;
    for (int m = 0; m < 9; ++m) v.mips[m] = mips ? mips + mip_offset(m) : nullptr;
    for (int m = 0; m < 8; ++m) v.tr[m] = tr ? tr + tr_offset(m) : nullptr;
    v.lut = lut;
    return v;
}

extern "C" {

// Environment map of every later call (SetHDRI / SetEnvExp); rgba NULL removes it.
void orc_set_env(const float* rgba, int w, int h, float env_exp) {
    g_env.clear();
    g_env_w = g_env_h = 0;
    g_env_exp = env_exp;
    if (rgba && w > 0 && h > 0) {
        g_env.assign(rgba, rgba + (size_t)w * h * 4);
        g_env_w = w;
        g_env_h = h;
    }
}

void orc_sky_box(const float* dir3, float* out3) {
    Volume v = to_volume(nullptr, 0, 1.0f, nullptr, nullptr, nullptr);
    f3 r = sky_box(v, mk(dir3[0], dir3[1], dir3[2]));
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}

int orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int orc_load_weights(const float* all, int total) {
    g_weights.assign(all, all + total);
    g_net_ok = net_bind(g_net, g_weights.data(), total);
    return g_net_ok ? 0 : 1;
}

void orc_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    mrp::u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    mrp::u32x4 r = mrp::philox4x32_10(c, k0, k1);
    memcpy(out, r.v, 16);
}

float orc_u01(uint32_t x) { return mrp::u01(x); }

// exp_f32 (fexp.h), the single-precision exp of the FP32 networks, on n arguments.
void orc_exp_f32(const float* x, int n, float* out) {
    for (int i = 0; i < n; ++i) out[i] = mrp::exp_f32(x[i]);
}

float orc_halton(unsigned dim, unsigned index) { return halton_scrambled(dim, index); }

void orc_stencil(int index, float* offset, int* mip, int* type, int* localindex) {
    StencilPoint s = get_samples23(index);
    *offset = s.offset;
    *mip = s.mip;
    *type = s.type;
    *localindex = s.localindex;
}

// feats: n rows of 577 = [X_Val | X_Val_Hg | X_Val_Sub | gamma]
int orc_predict_f32(const float* feats, int n, const float sr[3], float g, float* rgb) {
    if (!g_net_ok) return 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const float* f = feats + (size_t)i * NFEAT;
        predict_internal<float>(g_net, f, f + 2 * PC, f + PC, f[3 * PC], g, sr, rgb + 3 * (size_t)i);
    }
    return 0;
}
int orc_predict_f64(const double* feats, int n, const double sr[3], double g, double* rgb) {
    if (!g_net_ok) return 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const double* f = feats + (size_t)i * NFEAT;
        predict_internal<double>(g_net, f, f + 2 * PC, f + PC, f[3 * PC], g, sr, rgb + 3 * (size_t)i);
    }
    return 0;
}

// Per layer the largest |input| of the FP64 network over n rows (the FP8 calibration statistic).
int orc_predict_amax_f64(const double* feats, int n, const double sr[3], double g, double* amax) {
    if (!g_net_ok) return 1;
    for (int id = 0; id < kAmaxLayers; ++id) amax[id] = 0.0;
#pragma omp parallel
    {
        double loc[kAmaxLayers] = {};
        set_amax_recorder(&g_net, loc);
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            const double* f = feats + (size_t)i * NFEAT;
            double rgb[3];
            predict_internal<double>(g_net, f, f + 2 * PC, f + PC, f[3 * PC], g, sr, rgb);
        }
        set_amax_recorder(nullptr, nullptr);
#pragma omp critical
        for (int id = 0; id < kAmaxLayers; ++id) amax[id] = std::max(amax[id], loc[id]);
    }
    return 0;
}

void orc_build_pyramid(const float* grid, int N, float* mips_out) {
    float* m[9];
    for (int i = 0; i < 9; ++i) m[i] = mips_out + mip_offset(i);
    build_pyramid(grid, N, m);
}

void orc_bake_tr(const float* mips, const float light[3], float alpha, float* tr_out) {
    const float* m[9];
    for (int i = 0; i < 9; ++i) m[i] = mips + mip_offset(i);
    float* t[8];
    for (int i = 0; i < 8; ++i) t[i] = tr_out + tr_offset(i);
    bake_tr(m, mk(light[0], light[1], light[2]), alpha, t);
}

void orc_hg_lut_qmc(float g, uint32_t seed, float* lut) { hg_lut_qmc(g, seed, lut); }

double orc_hg_ball_mean(double costheta, double angle, double g, int n_psi, int n_phi) {
    return hg_ball_mean_fp64(costheta, angle, g, n_psi, n_phi);
}

float orc_sample_lut(const float* lut, float u, float v) { return sample_lut(lut, u, v); }

// Whole-image plain predictor.  out4: W*H*4; feats (optional): W*H*577; active/steps/pos optional.
int orc_render_predict(const float* grid, int N, float max_density, const float* mips, const float* tr,
                       const float* lut, const orc_view* ov, float* out4, float* feats, int32_t* active,
                       int32_t* steps, float* pos3) {
    if (!g_net_ok) return 1;
    Volume v = to_volume(grid, N, max_density, mips, tr, lut);
    View view = to_view(ov);
    int W = view.width, H = view.height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < W * H; ++p) {
        int row = p / W, col = p % W;
        int st = 0;
        predict_pixel(v, g_net, view, row, col, out4 + 4 * (size_t)p, feats ? feats + (size_t)p * NFEAT : nullptr,
                      &st);
        if (steps) steps[p] = st;
        if (active || pos3) {
            PixelState s = pixel_state(v, view, row, col);
            if (active) active[p] = s.active ? 1 : 0;
            if (pos3) { pos3[3 * p] = s.pos.x; pos3[3 * p + 1] = s.pos.y; pos3[3 * p + 2] = s.pos.z; }
        }
    }
    return 0;
}

void orc_rpnn_descriptor(const float* grid, int N, float max_density, const float* mips, const float* pos3,
                         const float* X3, const float* LX3, float alpha, float* desc, float* gamma) {
    Volume v = to_volume(grid, N, max_density, mips, nullptr, nullptr);
    rpnn_descriptor(v, mk(pos3[0], pos3[1], pos3[2]), mk(X3[0], X3[1], X3[2]), mk(LX3[0], LX3[1], LX3[2]), alpha,
                    desc, gamma);
}

// RenderCamera<true, RPNN> (volume.cu:198-228 -> render.cuh:808-875): per pixel the row of 2251
// floats the network reads (descriptor, gamma; zeros for inactive pixels), the active flag and
// the stored pixel max(0, (predict * light_color * scatter_rate, distance)), sky = SkyBox.
int orc_render_rpnn(const float* grid, int N, float max_density, const float* mips, const orc_view* ov,
                    float* rows, int32_t* active, float* out4) {
    if (!rpnn_ready()) return 1;
    Volume v = to_volume(grid, N, max_density, mips, nullptr, nullptr);
    View view = to_view(ov);
    const int W = view.width, H = view.height;
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < W * H; ++p) {
        PixelState st = pixel_state(v, view, p / W, p % W);
        float* row = rows + (size_t)p * 2251;
        float* o = out4 + 4 * (size_t)p;
        active[p] = st.active ? 1 : 0;
        if (!st.active) {
            memset(row, 0, 2251 * sizeof(float));
            const f3 sky = sky_box(v, st.dir);
            o[0] = fmaxf(0.0f, sky.x);
            o[1] = fmaxf(0.0f, sky.y);
            o[2] = fmaxf(0.0f, sky.z);
            o[3] = 0.0f;
            continue;
        }
        rpnn_descriptor(v, st.pos, st.XMain, st.LXMain, view.alpha, row, row + 2250);
        const float y = rpnn_predict_f32(row, row[2250]);
        const f3 r = (mk(y, y, y) * view.light_color) * view.scatter_rate;  // render.cuh:874
        o[0] = fmaxf(0.0f, r.x);
        o[1] = fmaxf(0.0f, r.y);
        o[2] = fmaxf(0.0f, r.z);
        o[3] = fmaxf(0.0f, length(view.ori - st.pos));
    }
    return 0;
}

// Tile feature gather (GatherData, volume.cu:630-708): W*H rows of 773 floats; active[p] = 1
// when some sub-ray of pixel p hits the medium.
int orc_tile_gather(const float* grid, int N, float max_density, const float* mips, const float* tr,
                    const float* lut, const orc_view* ov, int off_x, int off_y, float* out, int32_t* active) {
    Volume v = to_volume(grid, N, max_density, mips, tr, lut);
    View view = to_view(ov);
    int W = view.width, H = view.height;
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < W * H; ++p) {
        bool a = tile_gather_pixel(v, view, p / W, p % W, off_x, off_y, out + (size_t)p * TILE_ROW);
        if (active) active[p] = a ? 1 : 0;
    }
    return 0;
}
// One sub-ray's GetSampleT (render.cuh:464-486): ts = (t75, t25, first non-empty t, opacity).
void orc_sample_t(const float* grid, int N, float max_density, const float* ori3, const float* dir3, float alpha,
                  float* ts4) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    Ts4 t = get_sample_t(v, mk(ori3[0], ori3[1], ori3[2]), mk(dir3[0], dir3[1], dir3[2]), alpha);
    ts4[0] = t.x;
    ts4[1] = t.y;
    ts4[2] = t.z;
    ts4[3] = t.w;
}

// Time-bounded CPU baseline: renders the first n_pixels pixels in row-major order.
int orc_render_predict_rows(const float* grid, int N, float max_density, const float* mips, const float* tr,
                            const float* lut, const orc_view* ov, int p0, int n_pixels, float* out4) {
    if (!g_net_ok) return 1;
    Volume v = to_volume(grid, N, max_density, mips, tr, lut);
    View view = to_view(ov);
    int W = view.width;
#pragma omp parallel for schedule(dynamic, 16)
    for (int q = 0; q < n_pixels; ++q) {
        int p = p0 + q;
        predict_pixel(v, g_net, view, p / W, p % W, out4 + 4 * (size_t)q);
    }
    return 0;
}

// CPU baseline sample: the pixels listed in pix (row-major indices), OpenMP dynamic schedule;
// feats (optional): their 577 features.
int orc_render_predict_list(const float* grid, int N, float max_density, const float* mips, const float* tr,
                            const float* lut, const orc_view* ov, const int32_t* pix, int n, float* out4,
                            float* feats) {
    if (!g_net_ok) return 1;
    Volume v = to_volume(grid, N, max_density, mips, tr, lut);
    View view = to_view(ov);
    int W = view.width;
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < n; ++q)
        predict_pixel(v, g_net, view, pix[q] / W, pix[q] % W, out4 + 4 * (size_t)q,
                      feats ? feats + (size_t)q * NFEAT : nullptr);
    return 0;
}

// March only (GetSamplePoint) for arbitrary rays.
void orc_march(const float* grid, int N, float max_density, const float* ori3, const float* dir3, float alpha,
               float* pos3, int32_t* active, int32_t* taps) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    MarchResult m = get_sample_point(v, mk(ori3[0], ori3[1], ori3[2]), mk(dir3[0], dir3[1], dir3[2]), alpha);
    pos3[0] = m.pos.x;
    pos3[1] = m.pos.y;
    pos3[2] = m.pos.z;
    *active = m.active ? 1 : 0;
    *taps = m.steps;
}

// Features of one explicit state (pos, basis), for closed-form checks.
void orc_gather_state(const float* grid, int N, float max_density, const float* mips, const float* tr,
                      const float* lut, int active, const float* pos, const float* X, const float* Y,
                      const float* Z, const float* light, float g, float alpha, float* feats) {
    Volume v = to_volume(grid, N, max_density, mips, tr, lut);
    f3 L = mk(light[0], light[1], light[2]);
    gather_features(v, active != 0, mk(pos[0], pos[1], pos[2]), mk(X[0], X[1], X[2]), mk(Y[0], Y[1], Y[2]),
                    mk(Z[0], Z[1], Z[2]), normalize(L), L, g, alpha, feats);
}

void orc_basis(const float* dir, uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t view, float* out9) {
    Basis b = basis_from_normal(mk(dir[0], dir[1], dir[2]), seed, frame, pixel, view);
    float t[9] = {b.x.x, b.x.y, b.x.z, b.y.x, b.y.y, b.y.z, b.z.x, b.z.y, b.z.z};
    memcpy(out9, t, sizeof t);
}

int orc_pt_render(const float* grid, int N, float max_density, const float* tr, const orc_view* ov, int spp,
                  int max_bounce, int nee_baked, float tr_scale, double* mean3, double* m2_3, int64_t* taps_out) {
    Volume v = to_volume(grid, N, max_density, nullptr, tr, nullptr);
    View view = to_view(ov);
    PTParams pp{spp, max_bounce, nee_baked, tr_scale};
    int W = view.width, H = view.height;
    long long taps = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : taps)
    for (int p = 0; p < W * H; ++p) {
        long long t = 0;
        pt_pixel(v, view, pp, p / W, p % W, mean3 + 3 * (size_t)p, m2_3 + 3 * (size_t)p, &t);
        taps += t;
    }
    if (taps_out) *taps_out = taps;
    return 0;
}

// The path tracer on the listed pixels only (row-major indices): mean3 / m2_3 hold 3 doubles per
// listed pixel.
int orc_pt_render_list(const float* grid, int N, float max_density, const float* tr, const orc_view* ov,
                       const int32_t* pix, int n, int spp, int max_bounce, int nee_baked, float tr_scale,
                       double* mean3, double* m2_3) {
    Volume v = to_volume(grid, N, max_density, nullptr, tr, nullptr);
    View view = to_view(ov);
    PTParams pp{spp, max_bounce, nee_baked, tr_scale};
    const int W = view.width;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < n; ++q) {
        long long t = 0;
        pt_pixel(v, view, pp, pix[q] / W, pix[q] % W, mean3 + 3 * (size_t)q, m2_3 + 3 * (size_t)q, &t);
    }
    return 0;
}

// The path tracer with NEE from a cached transmittance bake (nee_baked = 2) on listed pixels.
int orc_pt_render_list_cached(const float* grid, int N, float max_density, const float* ctr, int ctr_res,
                              const orc_view* ov, const int32_t* pix, int n, int spp, int max_bounce, float tr_scale,
                              double* mean3, double* m2_3) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    View view = to_view(ov);
    PTParams pp{spp, max_bounce, 2, tr_scale, ctr, ctr_res};
    const int W = view.width;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < n; ++q) {
        long long t = 0;
        pt_pixel(v, view, pp, pix[q] / W, pix[q] % W, mean3 + 3 * (size_t)q, m2_3 + 3 * (size_t)q, &t);
    }
    return 0;
}

// Listed texels of the cached transmittance bake (Fill_BakeTR) of a res^3 grid.
void orc_bake_cached_tr(const float* grid, int N, float max_density, const float* light, float alpha, float tr_scale,
                        uint32_t seed, int res, const int64_t* texels, int n, float* out) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    const f3 L = normalize(mk(light[0], light[1], light[2]));
#pragma omp parallel for schedule(dynamic, 64)
    for (int q = 0; q < n; ++q) out[q] = cached_tr_texel(v, res, texels[q], L, alpha, tr_scale, seed);
}

// GetSamples: per task the FP64 mean / M2 over spp point-radiance paths (streams (seed, 0) / (i, s)).
void orc_point_radiance(const float* grid, int N, float max_density, int n, const float* ori3, const float* dir3,
                        const float* light3, const float* alpha, const float* g, const float* scatter,
                        const float* light_color, int multi_scatter, int spp, uint32_t seed, float tr_scale,
                        double* mean3, double* m2_3) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    const f3 lc = mk(light_color[0], light_color[1], light_color[2]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        const f3 ori = mk(ori3[3 * i], ori3[3 * i + 1], ori3[3 * i + 2]);
        const f3 dir = mk(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
        const f3 L = mk(light3[3 * i], light3[3 * i + 1], light3[3 * i + 2]);
        double mean[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
        for (int s = 0; s < spp; ++s) {
            mrp::PhiloxStream rs(seed, 0u, (uint32_t)i, mrp::RNG_POINT, (uint32_t)s);
            const f3 r = point_sample(v, rs, ori, dir, L, lc, scatter[i], alpha[i], g[i], multi_scatter, tr_scale);
            const double x[3] = {r.x, r.y, r.z};
            for (int c = 0; c < 3; ++c) {
                const double d = x[c] - mean[c];
                mean[c] += d / (double)(s + 1);
                m2[c] += d * (x[c] - mean[c]);
            }
        }
        for (int c = 0; c < 3; ++c) {
            mean3[3 * (size_t)i + c] = mean[c];
            m2_3[3 * (size_t)i + c] = m2[c];
        }
    }
}

// ShadowTerm_TR(s): per task the mean over spp of Tr(ori -> light) * HG(dot(dir, light), g).
void orc_tr_samples(const float* grid, int N, float max_density, int n, const float* ori3, const float* dir3,
                    const float* light, float alpha, float g, int spp, uint32_t seed, float tr_scale, double* out) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
    const f3 L = normalize(mk(light[0], light[1], light[2]));
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        const f3 ori = mk(ori3[3 * i], ori3[3 * i + 1], ori3[3 * i + 2]);
        const f3 dir = mk(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
        const float dis = ray_box_distance(ori, L);
        const float phase = henyey_greenstein(dot(dir, L), g);
        double acc = 0.0;
        for (int s = 0; s < spp; ++s) {
            mrp::PhiloxStream rs(seed, 2u, (uint32_t)i, mrp::RNG_TR, (uint32_t)s);
            acc += (double)(tr_stochastic(v, rs, ori, L, dis, alpha, tr_scale) * phase);
        }
        out[i] = acc / (double)spp;
    }
}

// Stochastic transmittance estimates (n independent streams) along one ray.
void orc_tr_stochastic(const float* grid, int N, float max_density, const float* pos, const float* dir, float dis,
                       float alpha, float tr_scale, uint32_t seed, int n, float* out) {
    Volume v = to_volume(grid, N, max_density, nullptr, nullptr, nullptr);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        mrp::PhiloxStream rs(seed, 0u, (uint32_t)i, mrp::RNG_PT, 0u);
        out[i] = tr_stochastic(v, rs, mk(pos[0], pos[1], pos[2]), mk(dir[0], dir[1], dir[2]), dis, alpha, tr_scale);
    }
}

// fBm volume generator (mrpnn_amd/csrc/fbm.h), same bits as the device generator.
void orc_fbm_volume(int N, int octaves, float base_freq, uint32_t seed, float* out) {
    int perm[512];
    mrp::fbm_make_perm(seed, perm);
    size_t n = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y)
            for (int z = 0; z < N; ++z)
                out[((size_t)x * N + y) * N + z] = mrp::fbm_voxel(perm, N, octaves, base_freq, x, y, z);
    float lo = out[0], hi = out[0];
    for (size_t i = 0; i < n; ++i) {
        lo = out[i] < lo ? out[i] : lo;
        hi = out[i] > hi ? out[i] : hi;
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) out[i] = mrp::fbm_density(out[i], lo, hi);
}

}  // extern "C"
