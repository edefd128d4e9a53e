// CPU oracle for the MRPNN per-pixel radiance-prediction path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared against; it is
// never linked into, loaded by or called from the product library.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// It restates, in plain C++, the arithmetic of the reference (sim00000n/MRPNN) for the path
// camera ray -> first-hit march -> basis -> 192-point feature gather -> MLP, plus the
// preparation steps (mip pyramid, transmittance mips, HG LUT) and the delta-tracking path
// tracer.  Every function cites the reference lines it follows.  fp32 code is compiled with
// -ffp-contract=off so float expressions round exactly as written; the MLP's
// multiply-accumulates use fmaf because nvcc contracts `Out += In * W` (fmad=true default).
//
// Deliberate, documented deviations from the reference (it cannot run here: CUDA only):
//   * texture filtering: exact fp32 trilinear/bilinear as the reference's host Sample()
//     (core/volume.cu:845-882) instead of the GPU's fixed-point texture weights;
//   * RNG: Philox4x32-10 streams (mrpnn_amd/csrc/philox.h) instead of cuRAND XORWOW;
//   * libm expf/logf/atanf/acosf/powf instead of CUDA -use_fast_math intrinsics;
//   * sky radiance is 0 (no environment map is bound in any workload);
//   * camera: i = row, j = column, optional aspect/FOV scale (1.0 reproduces the reference).
//
// Pinning status (what checks each function; see tests/test_oracle_*.py):
//   predict_internal  -- invariants (equal channels, output >= 0, inactive constant,
//                        permutation invariance of gates), FP32 vs FP64 agreement,
//                        independent numpy restatement of the dataflow.
//   gather            -- closed forms on constant/empty volumes, stencil geometry.
//   march             -- homogeneous-cube closed form, brute-force survival product.
//   pyramid           -- 2x box mean, FP64 sums, zero-border upsample brute force.
//   tr_bake           -- constant-density closed form away from borders.
//   hg_lut_qmc        -- FP64 quadrature of the same ball integral (statistical).
//   path tracer       -- empty volume, homogeneous transmittance, free-flight KS,
//                        furnace-style single-scatter closed form.
//   philox            -- Random123 known-answer vectors and rocRAND's philox header.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../mrpnn_amd/csrc/fexp.h"
#include "../mrpnn_amd/csrc/philox.h"
#include "../mrpnn_amd/csrc/ref_tables.h"

namespace oracle {

// ---------------------------------------------------------------------------------------
// float3 operators as the reference defines them (core/vector.cu:3-45)
// ---------------------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator*(f3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator/(f3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
inline f3 operator+(f3 a, float b) { return {a.x + b, a.y + b, a.z + b}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, float b) { return {a.x - b, a.y - b, a.z - b}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline f3 normalize(f3 v) {
    float len = sqrtf(dot(v, v));
    return {v.x / len, v.y / len, v.z / len};
}
inline float length(f3 a) { return sqrtf(dot(a, a)); }
inline float lerpf_(float a, float b, float v) { return a * (1.0f - v) + b * v; }  // vector.cu:33
inline f3 fmin3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
inline f3 fmax3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
// inv(): zero components map to 999999999 (vector.cu:19)
inline f3 inv(f3 a) {
    return {a.x == 0 ? 999999999.0f : 1.0f / a.x, a.y == 0 ? 999999999.0f : 1.0f / a.y,
            a.z == 0 ? 999999999.0f : 1.0f / a.z};
}

// RayBoxOffset: entry distance into [-0.4999,0.4999]^3 or -1 (vector.cu:46-63)
inline float ray_box_offset(f3 p, f3 dir) {
    dir = inv(dir);
    f3 bmax = {0.4999f, 0.4999f, 0.4999f};
    f3 to_axil = (-p) * dir;
    f3 axil_to_face = bmax * dir;
    f3 d0 = to_axil + axil_to_face, d1 = to_axil - axil_to_face;
    f3 tmin = fmin3(d0, d1), tmax = fmax3(d0, d1);
    float tmi = fmaxf(tmin.x, fmaxf(tmin.y, tmin.z));
    float tma = fminf(tmax.x, fminf(tmax.y, tmax.z));
    return tma >= tmi ? fmaxf(tmi, 0.0f) : -1.0f;
}
// RayBoxDistance: exit distance from [-0.5,0.5]^3 (vector.cu:65-80)
inline float ray_box_distance(f3 p, f3 dir) {
    dir = inv(dir);
    f3 bmax = {0.5f, 0.5f, 0.5f};
    f3 to_axil = (-p) * dir;
    f3 axil_to_face = bmax * dir;
    f3 d0 = to_axil + axil_to_face, d1 = to_axil - axil_to_face;
    f3 tmax = fmax3(d0, d1);
    return fminf(tmax.x, fminf(tmax.y, tmax.z));
}

// HenyeyGreenstein without 1/pi, scaled by 0.25 (vector.cu:210-214)
inline float henyey_greenstein(float cs, float g) {
    float g2 = g * g;
    return ((1.0f - g2) / powf((1.0f + g2 - 2.0f * g * cs), 1.5f)) * 0.25f;
}
// SampleHeneyGreenstein / SampleHenyeyGreenstein (vector.cu:216-242)
inline float sample_hg_cos(float s, float g) {
    if (fabsf(g) < 0.0001f) return s * 2.0f - 1.0f;
    float g2 = g * g;
    float t0 = (1.0f - g2) / (1.0f - g + 2.0f * g * s);
    return (1.0f + g2 - t0 * t0) / (2.0f * g);
}
inline f3 sample_hg(float e0, float e1, f3 v, float g) {
    float ct = sample_hg_cos(e0, g);
    const float phi = 6.28318531f * e1;  // 2 pi e1 (vector.cu:232), single precision as the library
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    f3 v2, v3;
    if ((double)fabsf(v.z) > sqrt(0.5)) v2 = normalize(cross(v, mk(0, 1, 0)));
    else v2 = normalize(cross(v, mk(0, 0, 1)));
    v3 = cross(v2, v);
    return (v3 * (st * cosf(phi))) + (v2 * (st * sinf(phi))) + (v * ct);
}
// UniformSampleSphere (vector.cu:195-208)
inline f3 uniform_sample_sphere(float ex, float ey) {
    float phi = 2.0f * 3.14159265359f * ex;
    float ct = 1.0f - 2.0f * ey;
    float st = sqrtf(1.0f - ct * ct);
    return {st * cosf(phi), st * sinf(phi), ct};
}

// ---------------------------------------------------------------------------------------
// Stencil schedule GetSamples23_ (vector.cu:304-405) and SphereRandom3 (vector.cu:265-302)
// ---------------------------------------------------------------------------------------
struct StencilPoint {
    float offset;    // cube units
    int mip;         // density mip; TR mip is max(mip-1,0)
    int type;        // 5 = sphere/ball point, 1 = point along the light direction
    int localindex;  // 0 only for the centre point (index 0)
};
inline StencilPoint get_samples23(int index) {
    const float ms = 0.6f;
    StencilPoint s;
    s.type = 5;
    s.localindex = 8;
    float off;
    if (index < 8) { off = (float)(1.0 / 256.0); off *= lerpf_(ms, 1.0f, 1.0f); s.mip = 0; s.localindex = index; }
    else if (index < 16) { off = (float)(2.5 / 256.0); off *= lerpf_(ms, 1.0f, 0.8f); s.mip = 1; }
    else if (index < 32) { off = (float)(5.5 / 256.0); off *= lerpf_(ms, 1.0f, 0.6f); s.mip = 2; }
    else if (index < 48) { off = (float)(11.5 / 256.0); off *= lerpf_(ms, 1.0f, 0.4f); s.mip = 3; }
    else if (index < 64) { off = (float)(23.5 / 256.0); off *= lerpf_(ms, 1.0f, 0.2f); s.mip = 4; }
    else if (index < 96) { off = (float)(47.5 / 256.0); off *= ms; s.mip = 5; }
    else if (index < 128) { off = (float)(95.5 / 256.0); off *= ms; s.mip = 6; }
    else if (index < 160) { off = (float)(191.5 / 256.0); off *= ms; s.mip = 7; }
    else {
        s.type = 1;
        if (index < 168) { off = (float)((index - 160) + 1); off = (float)(off / 256.0); s.mip = 0; }
        else if (index < 176) { off = (float)(8.5 + 1 + 2.0 * (index - 168)); off = (float)(off / 256.0); s.mip = 1; }
        else if (index < 184) { off = (float)(26.5 + 4.0 * (index - 176)); off = (float)(off / 256.0); s.mip = 2; }
        else { off = (float)(60.5 + 8.0 * (index - 184)); off = (float)(off / 256.0); s.mip = 3; }
    }
    s.offset = off;
    return s;
}
inline f3 stencil_dir(int index) {
    return {mrp::kStencilDirs[index][0], mrp::kStencilDirs[index][1], mrp::kStencilDirs[index][2]};
}

// ---------------------------------------------------------------------------------------
// Volume resources and texture semantics (renderResources.cu:32-102; volume.cu:845-882)
// ---------------------------------------------------------------------------------------
struct Volume {
    int N = 0;                 // full resolution
    float max_density = 1.0f;  // maxDensity symbol (volume.cu:1471)
    const float* grid = nullptr;           // N^3, (x*N+y)*N+z
    const float* mips[9] = {nullptr};      // (256>>m)^3
    const float* tr[8] = {nullptr};        // (128>>m)^3
    const float* lut = nullptr;            // 512*512, lut[j*512+i], i <-> cos, j <-> cone angle
    const float* env = nullptr;            // environment map, env_w*env_h RGBA, or null
    int env_w = 0, env_h = 0;
    float env_exp = 1.0f;
};

// SkyBox (render.cuh:24-27): tex2D of the environment map at (atan2(-z,x)/2pi + 0.5, acos(y)/pi),
// bilinear with wrap addressing (exact fp32 weights), times the exposure; black without a map.
inline f3 sky_box(const Volume& v, f3 dir) {
    if (!v.env) return f3{0.0f, 0.0f, 0.0f};
    const float u = atan2f(-dir.z, dir.x) * (float)(0.5 / 3.1415926) + 0.5f;
    const float t = acosf(fmaxf(fminf(dir.y, 1.0f), -1.0f)) * (float)(1.0 / 3.1415926);
    const float px = (u - floorf(u)) * (float)v.env_w - 0.5f, py = (t - floorf(t)) * (float)v.env_h - 0.5f;
    const float fx = floorf(px), fy = floorf(py);
    const float wx = px - fx, wy = py - fy;
    const int x0 = ((int)fx % v.env_w + v.env_w) % v.env_w, y0 = ((int)fy % v.env_h + v.env_h) % v.env_h;
    const int x1 = (x0 + 1) % v.env_w, y1 = (y0 + 1) % v.env_h;
    f3 r;
    float* o = &r.x;
    for (int c = 0; c < 3; ++c) {
        const float a = v.env[((size_t)y0 * v.env_w + x0) * 4 + c], b = v.env[((size_t)y0 * v.env_w + x1) * 4 + c];
        const float cc = v.env[((size_t)y1 * v.env_w + x0) * 4 + c], d = v.env[((size_t)y1 * v.env_w + x1) * 4 + c];
        o[c] = lerpf_(lerpf_(a, b, wx), lerpf_(cc, d, wx), wy);
    }
    return r * v.env_exp;
}

// Zero-border trilinear as the host Sample() (volume.cu:835-863), optional max(0,.) clamp.
inline float fetch_border(const float* d, int res, int x, int y, int z, bool clamp0) {
    if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return 0.0f;
    float v = d[((size_t)x * res + y) * res + z];
    return clamp0 ? fmaxf(0.0f, v) : v;
}
inline float sample_border(const float* d, int res, f3 uv, bool clamp0) {
    f3 pos = uv * (float)res - 0.5f;
    int x = (int)floorf(pos.x), y = (int)floorf(pos.y), z = (int)floorf(pos.z);
    f3 w = pos - mk((float)x, (float)y, (float)z);
    auto S = [&](int a, int b, int c) { return fetch_border(d, res, a, b, c, clamp0); };
    return lerpf_(lerpf_(lerpf_(S(x, y, z), S(x, y, z + 1), w.z), lerpf_(S(x, y + 1, z), S(x, y + 1, z + 1), w.z), w.y),
                  lerpf_(lerpf_(S(x + 1, y, z), S(x + 1, y, z + 1), w.z),
                         lerpf_(S(x + 1, y + 1, z), S(x + 1, y + 1, z + 1), w.z), w.y),
                  w.x);
}
// Clamp-addressing trilinear (SampleClamp, volume.cu:839-882): TR mips.
inline float fetch_clamp(const float* d, int res, int x, int y, int z) {
    x = x < 0 ? 0 : (x > res - 1 ? res - 1 : x);
    y = y < 0 ? 0 : (y > res - 1 ? res - 1 : y);
    z = z < 0 ? 0 : (z > res - 1 ? res - 1 : z);
    return fmaxf(0.0f, d[((size_t)x * res + y) * res + z]);
}
inline float sample_clamp(const float* d, int res, f3 uv) {
    f3 pos = uv * (float)res - 0.5f;
    int x = (int)floorf(pos.x), y = (int)floorf(pos.y), z = (int)floorf(pos.z);
    f3 w = pos - mk((float)x, (float)y, (float)z);
    auto S = [&](int a, int b, int c) { return fetch_clamp(d, res, a, b, c); };
    return lerpf_(lerpf_(lerpf_(S(x, y, z), S(x, y, z + 1), w.z), lerpf_(S(x, y + 1, z), S(x, y + 1, z + 1), w.z), w.y),
                  lerpf_(lerpf_(S(x + 1, y, z), S(x + 1, y, z + 1), w.z),
                         lerpf_(S(x + 1, y + 1, z), S(x + 1, y + 1, z + 1), w.z), w.y),
                  w.x);
}
// 512^2 HG LUT, normalised coords, linear filter, clamp (volume.cu:1413-1418).
constexpr int LUT = 512;
inline float sample_lut(const float* lut, float u, float v) {
    float px = u * (float)LUT - 0.5f, py = v * (float)LUT - 0.5f;
    int x = (int)floorf(px), y = (int)floorf(py);
    float wx = px - (float)x, wy = py - (float)y;
    auto C = [](int a) { return a < 0 ? 0 : (a > LUT - 1 ? LUT - 1 : a); };
    int x0 = C(x), x1 = C(x + 1), y0 = C(y), y1 = C(y + 1);
    float a = lerpf_(lut[y0 * LUT + x0], lut[y0 * LUT + x1], wx);
    float b = lerpf_(lut[y1 * LUT + x0], lut[y1 * LUT + x1], wx);
    return lerpf_(a, b, wy);
}

// Density(): full-res texture, border 0, raw values incl. negative SDF (render.cuh:34-38)
inline float density(const Volume& v, f3 pos) { return sample_border(v.grid, v.N, pos + 0.5f, false); }
// MipDensityDynamic (renderResources.cu:32-57)
inline float mip_density(const Volume& v, int mip, f3 pos) {
    if (mip < 0 || mip > 8) return 0.0f;
    return sample_border(v.mips[mip], 256 >> mip, pos + 0.5f, false);
}
// MipTrDynamic (renderResources.cu:59-82)
inline float mip_tr(const Volume& v, int mip, f3 pos) {
    if (mip < 0 || mip > 7) return 0.0f;
    return sample_clamp(v.tr[mip], 128 >> mip, pos + 0.5f);
}
// ShadowTerm_TRTex(...).x with lightColor 1 (renderResources.cu:84-102)
inline float shadow_term_tr(const Volume& v, f3 ori, f3 light_dir, int mip) {
    if ((double)ori.x < -0.5 || (double)ori.y < -0.5 || (double)ori.z < -0.5 || (double)ori.x > 0.5 ||
        (double)ori.y > 0.5 || (double)ori.z > 0.5) {
        float off = ray_box_offset(ori, light_dir);
        if (off >= 0) return mip_tr(v, mip, ori + light_dir * off);
        return 1.0f;
    }
    return mip_tr(v, mip, ori);
}

// ---------------------------------------------------------------------------------------
// First-hit march: DeterminateEstimatedHit bool version (render.cuh:105-197) and
// GetSamplePoint (render.cuh:436-462).  Returns active flag, hit position, step count.
// ---------------------------------------------------------------------------------------
struct MarchResult {
    bool active;
    f3 pos;
    int steps;  // density taps performed
};
inline bool determinate_estimated_hit(const Volume& v, float alpha, f3 pos, f3 dir, float dis, f3* next, int* taps) {
    float smax = v.max_density * alpha;
    float t = 0, last_non_empty_t = 0, prob = 1.0f;
    float t_step = 2.0f / (float)v.N;
    float pass_rate = expf(-t_step * smax);
    int loop = 0;
    float last_dens = 0;
    while (loop++ < 10000) {
        if (t > dis) {
            if (last_non_empty_t != 0.0f) { t = last_non_empty_t; break; }
            *next = mk(0, 0, 0);
            return false;
        }
        float d = density(v, pos + (dir * t));
        ++*taps;
        if (d > 0) {
            float S = d * alpha;
            float sp = pass_rate + (1.0f - pass_rate) * (1.0f - S / smax);
            float np = prob * sp;
            if (np < 0.5f) {
                float off = 1.0f - logf(0.5f / prob) / logf(sp);
                t -= t_step * off;
                break;
            }
            prob = np;
            t += t_step;
        } else {
            if (last_dens > 0) {
                float S = last_dens * alpha;
                float ts = t_step + d;
                float pr = expf(-(ts * smax));
                float sp = pr + (1.0f - pr) * (1.0f - S / smax);
                float np = prob * sp;
                if (np < 0.5f) {
                    float off = 1.0f - logf(0.5f / prob) / logf(sp);
                    t -= ts * off;
                    break;
                }
                prob = np;
                last_non_empty_t = t + d;
            }
            t += fmaxf(-d * 1.0f, 1.0f / (float)v.N);
        }
        last_dens = d;
    }
    *next = (dir * t) + pos;
    return true;
}
inline MarchResult get_sample_point(const Volume& v, f3 ori, f3 dir, float alpha) {
    MarchResult r{false, mk(0, 0, 0), 0};
    dir = normalize(dir);
    float dis = ray_box_offset(ori, dir);
    if (dis < 0) return r;
    f3 sp = ori + dir * dis;
    float maxd = ray_box_distance(sp, dir);
    f3 p0;
    bool in = determinate_estimated_hit(v, alpha, sp, dir, maxd, &p0, &r.steps);
    if (in) { r.active = true; r.pos = p0; }
    return r;
}

// ---------------------------------------------------------------------------------------
// Basis GetMatrixFromNormal(seed, v1, 0.1) (render.cuh:694-708) with a Philox stream.
// Attempts are capped at 64 (the reference loops until |dot(r,v1)| >= 1e-5).
// ---------------------------------------------------------------------------------------
struct Basis {
    f3 x, y, z;
};
inline Basis basis_from_normal(f3 v1, uint32_t seed, uint32_t frame, uint32_t pixel, uint32_t view) {
    const float spread = 0.1f;  // render.cuh:822
    v1 = normalize(v1);
    mrp::PhiloxStream rs(seed, frame, pixel, mrp::RNG_BASIS_MAIN, view);
    f3 v2 = mk(0, 1, 0), v3 = mk(0, 0, 1);
    for (int att = 0; att < 64; ++att) {
        float e0 = rs.next() * spread;
        float e1 = rs.next() * spread;
        f3 r = uniform_sample_sphere(e0, e1);
        if ((double)fabsf(dot(r, v1)) < 0.00001) continue;
        v2 = normalize(cross(v1, r));
        v3 = cross(v2, v1);
        break;
    }
    return {v1, v2, v3};
}

// ---------------------------------------------------------------------------------------
// Feature gather SetupInputData (radiancePredict.cu:427-485).  Row layout of the 577
// features: [X_Val 0:192 | X_Val_Hg 192:384 | X_Val_Sub 384:576 | gamma].
// ---------------------------------------------------------------------------------------
constexpr int PC = 192;
constexpr int NFEAT = 577;
inline void gather_features(const Volume& v, bool active, f3 pos, f3 XMain, f3 YMain, f3 ZMain, f3 LXMain,
                            f3 LightDir, float g, float alpha, float* out) {
    float* Xv = out;
    float* Xh = out + PC;
    float* Xs = out + 2 * PC;
    if (active) {
        for (int i = 0; i < PC; ++i) {
            StencilPoint s = get_samples23(i);
            float layer = (float)s.mip + 0.1f;
            int mip = (int)layer;
            f3 p;
            if (s.type >= 4) {
                f3 L = stencil_dir(i);
                p = pos + ((XMain * L.x + YMain * L.y + ZMain * L.z) * s.offset);
            } else {
                p = pos + LightDir * s.offset;
            }
            Xv[i] = logf(1.0f + alpha * mip_density(v, mip, p) / 64.0f);
            if (s.localindex == 0) {
                Xh[i] = logf(henyey_greenstein(dot(XMain, LightDir), g) + 1.0f);
            } else {
                f3 ms = normalize(p - pos);
                float radius = (float)(1 << mip) / 256.0f;
                float angle = atanf(0.5f * radius / s.offset);
                float c1 = dot(XMain, ms), c2 = dot(LXMain, ms);
                float u = c1 * 0.5f + 0.5f, u2 = c2 * 0.5f + 0.5f;
                float vv = angle / (3.1415926535f * 60.0f / 180.0f);
                float h0 = sample_lut(v.lut, u, vv), h1 = sample_lut(v.lut, u2, vv);
                Xh[i] = logf(h0 * h1 + 1.0f);
            }
            Xs[i] = shadow_term_tr(v, p, LXMain, mip - 1 > 0 ? mip - 1 : 0);
        }
    } else {
        for (int i = 0; i < PC; ++i) Xv[i] = Xh[i] = Xs[i] = 0.0f;
    }
    float d = dot(XMain, LXMain);
    d = fminf(1.0f, fmaxf(-1.0f, d));
    out[3 * PC] = acosf(d);
}

// ---------------------------------------------------------------------------------------
// Network weights (core/NNWeight.cuh, header order) and RadiancePredictInternal
// (radiancePredict.cu:11-422).
// ---------------------------------------------------------------------------------------
struct Net {
    const float* p[146];
    const float* get(const char* name) const;
};
bool net_bind(Net& net, const float* all, int total);

template <class T> inline T mac(T acc, T a, T b);
template <> inline float mac<float>(float acc, float a, float b) { return fmaf(a, b, acc); }
template <> inline double mac<double>(double acc, double a, double b) { return acc + a * b; }

template <class T>
void predict_internal(const Net& net, const T* X_Val, const T* X_Val_Sub, const T* X_Val_Hg, T gamma, T g,
                      const T scatterrate[3], T out[3]);
// While set (per thread; net = nullptr switches it off), predict_internal folds the largest
// |input| of each of the 60 matrix-product layers into amax[id]: L01..L81 of the three branches
// (L, L_Tr, L_Hg), LD01..LD31 of the three, LD41, LD_Tr41, LD_Hg41, LSEFin2, LC0, LC1, LC2, LX,
// LX0..LX12 -- the FP8 layers of the library's C ABI in its order.
constexpr int kAmaxLayers = 60;
void set_amax_recorder(const Net* net, double* amax);

// One pixel of the plain per-frame predictor: RenderCamera<true,MRPNN> (volume.cu:198-228)
// -> NNPredict<MRPNN> (render.cuh:808-875).  out = (rgb, distance) after max(0,.).
struct View {
    f3 ori, up, right, light_dir, light_color, scatter_rate;
    float alpha, g, tan_half_fov;
    int width, height;
    uint32_t seed, frame, view_index;
    float jitter;  // < 0: Philox jitter per pixel; else fixed subpixel offset
};
struct PixelState {
    bool active;
    f3 pos, XMain, YMain, ZMain, LXMain, LightDir, dir;
    int steps;
};
PixelState pixel_state(const Volume& v, const View& view, int row, int col);
void predict_pixel(const Volume& v, const Net& net, const View& view, int row, int col, float out4[4],
                   float* feats_out = nullptr, int* steps_out = nullptr);

// ---------------------------------------------------------------------------------------
// Tile feature gather (training-data path): GatherSamples (volume.cu:257-283) ->
// NNGatherSamples (render.cuh:1049-1142) -> PrepareRay (render.cuh:1006-1047) -> GetSampleT
// (render.cuh:464-486) -> DeterminateEstimatedHit float4 version with CheckExitCondition
// (render.cuh:199-340) -> PopulateInputData / SetupInputData2 (radiancePredict.cu:487-565).
// One output row of 773 floats per pixel: [X_Val 0:192 | X_Val_Hg 192:384 | X_Val_Sub 384:576 |
// gamma 576:581 | t grid 581:773] (mrpnn_helpers.py unpack_features).
// ---------------------------------------------------------------------------------------
constexpr int TILE_ROW = 773;
// -logf(0.75f), -logf(0.76f), -logf(0.25f), correctly rounded (the reference evaluates logf)
constexpr float kTauFirst = 0x1.269622p-2f, kTauStepMax = 0x1.1905fap-2f, kTauSecond = 0x1.62e43p+0f;
struct Ts4 {
    float x, y, z, w;  // t at survival 0.75, t at survival 0.25, first non-empty t, opacity
};
Ts4 estimated_hit_tau(const Volume& v, float alpha, f3 pos, f3 dir, float dis);
Ts4 get_sample_t(const Volume& v, f3 ori, f3 dir, float alpha);
// out: 773 floats.  Returns false for a pixel none of whose 64 sub-rays hits the medium (the
// reference then writes only out[0] = -1000 and the t grid; here the rest of the row is 0).
bool tile_gather_pixel(const Volume& v, const View& view, int row, int col, int off_x, int off_y, float* out);

// RPNN comparison predictor (rpnn.cpp): RadiancePredict_RPNN (radiancePredict.cu:567-722)
void rpnn_descriptor(const Volume& v, f3 pos, f3 X, f3 LX, float alpha, float* desc /* 2250 */, float* gamma);
bool rpnn_ready();
float rpnn_predict_f32(const float* desc, float gamma);

// Preparation steps
void build_pyramid(const float* grid, int N, float* const mips[9]);                 // volume.cu:1474-1515
void bake_tr(const float* const mips[9], f3 light_dir, float alpha, float* const tr[8]);  // volume.cu:1535-1560,1636-1653
void hg_lut_qmc(float g, uint32_t seed, float* lut);                                 // volume.cu:1322-1400
double hg_ball_mean_fp64(double costheta, double angle, double g, int n_psi, int n_phi);
float halton_scrambled(unsigned dim, unsigned index);                                // volume.cu:1300-1319

// Path tracer (render.cuh:60-103, 399-434, 535-597) with Philox and FP64 Welford.
struct PTParams {
    int spp, max_bounce;
    int nee_baked;  // 0: stochastic Tr; 1: TR mip 0 (render.cuh:589); 2: the cached bake below
    float tr_scale;
    const float* ctr = nullptr;  // nee_baked 2: ctr_res^3 cached transmittance texels
    int ctr_res = 0;
};
float tr_stochastic(const Volume& v, mrp::PhiloxStream& rs, f3 pos, f3 dir, float dis, float alpha, float tr_scale);
void pt_pixel(const Volume& v, const View& view, const PTParams& pp, int row, int col, double* mean3, double* m2_3,
              long long* taps);
// Fill_BakeTR (volume.cu:1562-1587) texel i of a res^3 grid: mean of 5 (density > 0 at the
// centre) or 2 stochastic Tr estimates towards light (unit), estimate e from the Philox stream
// (seed, 1) / (i, RNG_BAKE, e).
float cached_tr_texel(const Volume& v, int res, long long i, f3 light, float alpha, float tr_scale, uint32_t seed);
// GetSample (render.cuh:615-654), one path from stream rs.
f3 point_sample(const Volume& v, mrp::PhiloxStream& rs, f3 ori, f3 dir, f3 light, f3 light_color, float scatter,
                float alpha, float g, int max_bounce, float tr_scale);

}  // namespace oracle
