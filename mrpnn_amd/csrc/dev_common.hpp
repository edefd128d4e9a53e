// Device building blocks of the MRPNN predictor path (gfx950).
//
// Compiled with -ffp-contract=off: every float expression rounds as written, and fused
// multiply-adds appear only where written as fmaf().  The per-pixel functions here follow
// the reference line by line (citations below) so the plain kernel reproduces the CPU
// restatement up to transcendental-function ulps.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "philox.h"

namespace mrp {

// ---------------------------------------------------------------------------------------
// float3 arithmetic with the reference's operator set (core/vector.cu:3-45)
// ---------------------------------------------------------------------------------------
struct v3 {
    float x, y, z;
};
MRP_HD v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
MRP_HD v3 operator*(v3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
MRP_HD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
MRP_HD v3 operator/(v3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
MRP_HD v3 operator+(v3 a, float b) { return {a.x + b, a.y + b, a.z + b}; }
MRP_HD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MRP_HD v3 operator-(v3 a, float b) { return {a.x - b, a.y - b, a.z - b}; }
MRP_HD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MRP_HD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
MRP_HD float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MRP_HD v3 cross3(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
MRP_HD v3 normalize3(v3 v) {
    float len = sqrtf(dot3(v, v));
    return {v.x / len, v.y / len, v.z / len};
}
MRP_HD float length3(v3 a) { return sqrtf(dot3(a, a)); }
MRP_HD float lerp_ref(float a, float b, float v) { return a * (1.0f - v) + b * v; }  // vector.cu:33
MRP_HD v3 vmin3(v3 a, v3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
MRP_HD v3 vmax3(v3 a, v3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
MRP_HD v3 inv3(v3 a) {  // vector.cu:19
    return {a.x == 0 ? 999999999.0f : 1.0f / a.x, a.y == 0 ? 999999999.0f : 1.0f / a.y,
            a.z == 0 ? 999999999.0f : 1.0f / a.z};
}

// RayBoxOffset (vector.cu:46-63): entry distance into [-0.4999,0.4999]^3, -1 on a miss.
MRP_HD float ray_box_offset(v3 p, v3 dir) {
    dir = inv3(dir);
    v3 bmax = {0.4999f, 0.4999f, 0.4999f};
    v3 to_axil = (-p) * dir;
    v3 a2f = bmax * dir;
    v3 d0 = to_axil + a2f, d1 = to_axil - a2f;
    v3 tmin = vmin3(d0, d1), tmax = vmax3(d0, d1);
    float tmi = fmaxf(tmin.x, fmaxf(tmin.y, tmin.z));
    float tma = fminf(tmax.x, fminf(tmax.y, tmax.z));
    return tma >= tmi ? fmaxf(tmi, 0.0f) : -1.0f;
}
// RayBoxDistance (vector.cu:65-80): exit distance from [-0.5,0.5]^3.
MRP_HD float ray_box_distance(v3 p, v3 dir) {
    dir = inv3(dir);
    v3 bmax = {0.5f, 0.5f, 0.5f};
    v3 to_axil = (-p) * dir;
    v3 a2f = bmax * dir;
    v3 d0 = to_axil + a2f, d1 = to_axil - a2f;
    v3 tmax = vmax3(d0, d1);
    return fminf(tmax.x, fminf(tmax.y, tmax.z));
}

// HenyeyGreenstein * pi / (4 pi) convention: no 1/pi, x0.25 (vector.cu:210-214)
MRP_HD float henyey_greenstein(float cs, float g) {
    float g2 = g * g;
    return ((1.0f - g2) / powf((1.0f + g2 - 2.0f * g * cs), 1.5f)) * 0.25f;
}
MRP_HD float sample_hg_cos(float s, float g) {  // vector.cu:216-223
    if (fabsf(g) < 0.0001f) return s * 2.0f - 1.0f;
    float g2 = g * g;
    float t0 = (1.0f - g2) / (1.0f - g + 2.0f * g * s);
    return (1.0f + g2 - t0 * t0) / (2.0f * g);
}
MRP_HD v3 sample_hg(float e0, float e1, v3 v, float g) {  // vector.cu:225-242
    float ct = sample_hg_cos(e0, g);
    const float phi = 6.28318531f * e1;  // 2 pi e1 (vector.cu:232, double there)
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    v3 v2;
    if (fabsf(v.z) > 0.70710678118654757f) v2 = normalize3(cross3(v, mk3(0, 1, 0)));
    else v2 = normalize3(cross3(v, mk3(0, 0, 1)));
    v3 v3_ = cross3(v2, v);
    return (v3_ * (st * cosf(phi))) + (v2 * (st * sinf(phi))) + (v * ct);
}
MRP_HD v3 uniform_sample_sphere(float ex, float ey) {  // vector.cu:195-208
    float phi = 2.0f * 3.14159265359f * ex;
    float ct = 1.0f - 2.0f * ey;
    float st = sqrtf(1.0f - ct * ct);
    return {st * cosf(phi), st * sinf(phi), ct};
}

// ---------------------------------------------------------------------------------------
// Resident volume: full-res grid, density mips, TR mips, HG LUT (all fp32, HBM)
// ---------------------------------------------------------------------------------------
struct DevVolume {
    const float* grid;     // N^3, (x*N+y)*N+z
    const float* mips[9];  // (256>>m)^3
    const float* tr[8];    // (128>>m)^3
    const float* lut;      // 512*512, [j*512+i]
    const float* ptr;      // (y, z) quads of the zero-padded TR mips 0..6, 4 floats each (fast gather)
    const float4* qmips;   // (y, z) difference quads of the zero-padded density mips 0..7 (fast gather)
    const uint8_t* occ;    // brick emptiness of the grid, occ_n^3 (see brick_of)
    int occ_n;
    int N;
    float max_density;
    const float4* env;  // environment map, env_w * env_h RGBA texels (row-major), or null
    int env_w, env_h;
    float env_exp;      // enviroment_exp (core/render.cuh:15)
};

// SkyBox (core/render.cuh:24-27): the environment map at (atan2(-z, x) / 2pi + 0.5,
// acos(y) / pi) with bilinear filtering and wrap addressing, times the exposure; zero radiance
// without a map (the reference's unbound texture).
__device__ inline v3 sky_box(const DevVolume& v, v3 dir) {
    if (!v.env) return {0.0f, 0.0f, 0.0f};
    const float u = atan2f(-dir.z, dir.x) * (float)(0.5 / 3.1415926) + 0.5f;
    const float t = acosf(fmaxf(fminf(dir.y, 1.0f), -1.0f)) * (float)(1.0 / 3.1415926);
    const float px = (u - floorf(u)) * (float)v.env_w - 0.5f, py = (t - floorf(t)) * (float)v.env_h - 0.5f;
    const float fx = floorf(px), fy = floorf(py);
    const float wx = px - fx, wy = py - fy;
    int x0 = (int)fx, y0 = (int)fy;
    x0 = x0 < 0 ? x0 + v.env_w : x0;
    y0 = y0 < 0 ? y0 + v.env_h : y0;
    const int x1 = x0 + 1 == v.env_w ? 0 : x0 + 1, y1 = y0 + 1 == v.env_h ? 0 : y0 + 1;
    const float4 a = v.env[(size_t)y0 * v.env_w + x0], b = v.env[(size_t)y0 * v.env_w + x1];
    const float4 c = v.env[(size_t)y1 * v.env_w + x0], d = v.env[(size_t)y1 * v.env_w + x1];
    auto bl = [&](float p, float q, float r, float s) {
        return lerp_ref(lerp_ref(p, q, wx), lerp_ref(r, s, wx), wy);
    };
    return mk3(bl(a.x, b.x, c.x, d.x), bl(a.y, b.y, c.y, d.y), bl(a.z, b.z, c.z, d.z)) * v.env_exp;
}

// Stencil entry, host-precomputed from GetSamples23_ (vector.cu:304-405) and the
// direction tables (vector.cu:82-177).
struct StencilEntry {
    float dx, dy, dz;  // direction in (XMain,YMain,ZMain) frame (type 5)
    float offset;      // cube units
    float lut_v;       // LUT v = atan(0.5*2^mip/256/offset) / (pi/3)  (radiancePredict.cu:458-464)
    int mip;           // density mip
    int type;          // 5 sphere/ball, 1 light ray
    int local;         // 0 = centre point
};
constexpr int kPC = 192;
constexpr int kNFEAT = 577;

// Exact zero-border trilinear as the host Sample() (core/volume.cu:835-863); raw values.
__device__ inline float fetch_b(const float* __restrict__ d, int res, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return 0.0f;
    return d[((size_t)x * res + y) * res + z];
}
__device__ inline float sample_border_exact(const float* __restrict__ d, int res, v3 uv) {
    v3 pos = uv * (float)res - 0.5f;
    int x = (int)floorf(pos.x), y = (int)floorf(pos.y), z = (int)floorf(pos.z);
    v3 w = pos - mk3((float)x, (float)y, (float)z);
    return lerp_ref(lerp_ref(lerp_ref(fetch_b(d, res, x, y, z), fetch_b(d, res, x, y, z + 1), w.z),
                             lerp_ref(fetch_b(d, res, x, y + 1, z), fetch_b(d, res, x, y + 1, z + 1), w.z), w.y),
                    lerp_ref(lerp_ref(fetch_b(d, res, x + 1, y, z), fetch_b(d, res, x + 1, y, z + 1), w.z),
                             lerp_ref(fetch_b(d, res, x + 1, y + 1, z), fetch_b(d, res, x + 1, y + 1, z + 1), w.z),
                             w.y),
                    w.x);
}
// Clamp addressing (SampleClamp, volume.cu:839-844,864-882): TR mips.
__device__ inline float fetch_c(const float* __restrict__ d, int res, int x, int y, int z) {
    x = x < 0 ? 0 : (x > res - 1 ? res - 1 : x);
    y = y < 0 ? 0 : (y > res - 1 ? res - 1 : y);
    z = z < 0 ? 0 : (z > res - 1 ? res - 1 : z);
    return fmaxf(0.0f, d[((size_t)x * res + y) * res + z]);
}
__device__ inline float sample_clamp_exact(const float* __restrict__ d, int res, v3 uv) {
    v3 pos = uv * (float)res - 0.5f;
    int x = (int)floorf(pos.x), y = (int)floorf(pos.y), z = (int)floorf(pos.z);
    v3 w = pos - mk3((float)x, (float)y, (float)z);
    return lerp_ref(lerp_ref(lerp_ref(fetch_c(d, res, x, y, z), fetch_c(d, res, x, y, z + 1), w.z),
                             lerp_ref(fetch_c(d, res, x, y + 1, z), fetch_c(d, res, x, y + 1, z + 1), w.z), w.y),
                    lerp_ref(lerp_ref(fetch_c(d, res, x + 1, y, z), fetch_c(d, res, x + 1, y, z + 1), w.z),
                             lerp_ref(fetch_c(d, res, x + 1, y + 1, z), fetch_c(d, res, x + 1, y + 1, z + 1), w.z),
                             w.y),
                    w.x);
}
// HG LUT: 512^2, normalised coordinates, linear filter, clamp (volume.cu:1413-1418).
__device__ inline float sample_lut_exact(const float* __restrict__ lut, float u, float v) {
    float px = u * 512.0f - 0.5f, py = v * 512.0f - 0.5f;
    int x = (int)floorf(px), y = (int)floorf(py);
    float wx = px - (float)x, wy = py - (float)y;
    int x0 = min(max(x, 0), 511), x1 = min(max(x + 1, 0), 511);
    int y0 = min(max(y, 0), 511), y1 = min(max(y + 1, 0), 511);
    float a = lerp_ref(lut[y0 * 512 + x0], lut[y0 * 512 + x1], wx);
    float b = lerp_ref(lut[y1 * 512 + x0], lut[y1 * 512 + x1], wx);
    return lerp_ref(a, b, wy);
}

// Full-resolution zero-border trilinear with the value of sample_border_exact(v.grid, N, p+0.5),
// bit for bit (same coordinates, weights and lerp order), fetched as four (z, z+1) pairs:
// 64-bit global loads at clamped in-grid indices, one 64-bit slab address per x tap and
// 32-bit offsets inside a slab, then the border taps are zeroed by selects (no branches).
// Needs N >= 2.
// a * b + c for a, b < 2^24 (v_mad_u32_u24, full rate).  Written out: for the texel index the
// compiler otherwise emits a 64-bit multiply-add or a 32-bit multiply (both quarter rate)
// although every operand is a small non-negative integer.  _s: b wave-uniform (scalar).
__device__ inline unsigned mad_u24_s(unsigned a, unsigned b, unsigned c) {
    unsigned d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}
__device__ inline unsigned mad_u24_v(unsigned a, unsigned b, unsigned c) {
    unsigned d;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

typedef float f32x2u_t __attribute__((ext_vector_type(2), aligned(4)));
typedef __attribute__((address_space(1))) const f32x2u_t* gpair_ptr;
struct FullTap {  // integer corner and weights of one sample
    int x, y, z;
    v3 w;
};
__device__ inline FullTap full_tap(int N, v3 p) {
    const v3 uv = p + 0.5f;
    const v3 pos = uv * (float)N - 0.5f;
    FullTap t;
    const v3 fl = mk3(floorf(pos.x), floorf(pos.y), floorf(pos.z));
    t.x = (int)fl.x;
    t.y = (int)fl.y;
    t.z = (int)fl.z;
    t.w = pos - fl;
    return t;
}
// All eight taps inside the grid (corner in [0, N-2]^3): pair loads at 32-bit byte offsets
// ((x*N + y)*N + z)*4 from a buffer resource, no border selects; the same taps and blend as
// the border path below.  Grids up to 1023^3 (byte offsets below 2^32).
__device__ inline bool full_tap_inner(int N, const FullTap& t) {
    return max((unsigned)t.x, max((unsigned)t.y, (unsigned)t.z)) <= (unsigned)(N - 2);
}
__device__ inline float density_inner(const DevVolume& v, const FullTap& t) {
    const unsigned N = (unsigned)v.N;
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)v.grid, (short)0, (int)(N * N * N * 4u), 0x00020000);
    const unsigned o = mad_u24_v(mad_u24_v((unsigned)t.x, N, (unsigned)t.y), N, (unsigned)t.z) * 4u;
    const unsigned s1 = 4u * N, s2 = 4u * N * N;
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 c00 = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o, 0, 0));
    const f2 c01 = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(o + s1), 0, 0));
    const f2 c10 = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(o + s2), 0, 0));
    const f2 c11 = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(o + s2 + s1), 0, 0));
    const v3 w = t.w;
    return lerp_ref(lerp_ref(lerp_ref(c00.x, c00.y, w.z), lerp_ref(c01.x, c01.y, w.z), w.y),
                    lerp_ref(lerp_ref(c10.x, c10.y, w.z), lerp_ref(c11.x, c11.y, w.z), w.y), w.x);
}
__device__ inline float density_border_at(const DevVolume& v, const FullTap& t);
// Interior fast path when no lane of the wavefront is near a face, else the border path.
__device__ inline float density_full_at(const DevVolume& v, const FullTap& t) {
    if (v.N <= 1023 && __builtin_amdgcn_ballot_w64(!full_tap_inner(v.N, t)) == 0) return density_inner(v, t);
    return density_border_at(v, t);
}
__device__ inline float density_border_at(const DevVolume& v, const FullTap& t) {
    const int N = v.N;
    const int z0 = min(max(t.z, 0), N - 2);  // pair (z0, z0 + 1) lies inside the row
    const bool zlo = t.z >= 0, zhi = t.z + 1 < N;
    // t.z = -1: wanted (0, v[0]) = (0, pair.x); t.z = N-1: wanted (v[N-1], 0) = (pair.y, 0)
    const bool shift = t.z == N - 1;
    float c[2][2][2];
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
        const int xi = t.x + dx;
        const bool xok = (unsigned)xi < (unsigned)N;
        const float* slab = v.grid + (size_t)min(max(xi, 0), N - 1) * N * N;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int yi = t.y + dy;
            const bool ok = xok && (unsigned)yi < (unsigned)N;
            const int off = min(max(yi, 0), N - 1) * N + z0;
            const f32x2u_t pr = *(gpair_ptr)(slab + off);
            const float a = shift ? pr.y : pr.x;
            c[dx][dy][0] = (ok && zlo) ? a : 0.0f;
            c[dx][dy][1] = (ok && zhi) ? (t.z < 0 ? pr.x : pr.y) : 0.0f;
        }
    }
    const v3 w = t.w;
    return lerp_ref(lerp_ref(lerp_ref(c[0][0][0], c[0][0][1], w.z), lerp_ref(c[0][1][0], c[0][1][1], w.z), w.y),
                    lerp_ref(lerp_ref(c[1][0][0], c[1][0][1], w.z), lerp_ref(c[1][1][0], c[1][1][1], w.z), w.y),
                    w.x);
}

// Density() full-res border (render.cuh:34-38); MipDensityDynamic (renderResources.cu:32-57);
// MipTrDynamic (:59-82); ShadowTerm_TRTex(...).x (:84-102).
__device__ inline float density_full(const DevVolume& v, v3 p) { return density_full_at(v, full_tap(v.N, p)); }
__device__ inline float mip_density(const DevVolume& v, int mip, v3 p) {
    return sample_border_exact(v.mips[mip], 256 >> mip, p + 0.5f);
}
__device__ inline float mip_tr(const DevVolume& v, int mip, v3 p) {
    return sample_clamp_exact(v.tr[mip], 128 >> mip, p + 0.5f);
}
__device__ inline float shadow_term_tr(const DevVolume& v, v3 ori, v3 light, int mip) {
    if (ori.x < -0.5f || ori.y < -0.5f || ori.z < -0.5f || ori.x > 0.5f || ori.y > 0.5f || ori.z > 0.5f) {
        float off = ray_box_offset(ori, light);
        if (off >= 0) return mip_tr(v, mip, ori + light * off);
        return 1.0f;
    }
    return mip_tr(v, mip, ori);
}

// ---------------------------------------------------------------------------------------
// First-hit march: DeterminateEstimatedHit, bool version (render.cuh:105-197), and
// GetSamplePoint (render.cuh:436-462).  taps counts full-res density taps.
// ---------------------------------------------------------------------------------------
// Empty-brick test of the first-hit march.  The grid is covered by bricks of 8 texel
// corners: a sample whose lower trilinear corner is x0 (per axis) lies in brick (x0+1)>>3 and
// reads texels x0, x0+1 in [8b-1, 8b+7].  occ[b] = 1 when every texel of that range (inside
// the grid) is exactly zero, so such a sample is exactly 0 and its fetch can be skipped
// without changing any result.
// The march keeps the last brick it looked up (key, emptiness) in registers: a step stays in
// one brick for several samples, so most steps need no occupancy load.
struct BrickCache {
    int key = -2;  // -1: outside the brick grid
    bool empty = false;
};
__device__ inline bool tap_in_empty_brick(const DevVolume& v, const FullTap& t, BrickCache& bc) {
    const int n = v.occ_n;
    const int bx = (t.x + 1) >> 3, by = (t.y + 1) >> 3, bz = (t.z + 1) >> 3;
    const bool in = (unsigned)bx < (unsigned)n && (unsigned)by < (unsigned)n && (unsigned)bz < (unsigned)n;
    const int key = in ? (int)mad_u24_v(mad_u24_v((unsigned)bx, (unsigned)n, (unsigned)by), (unsigned)n, (unsigned)bz) : -1;
    if (key != bc.key) {
        bc.key = key;
        bc.empty = in && *(const __attribute__((address_space(1))) uint8_t*)(v.occ + key) != 0;
    }
    return bc.empty;
}

__device__ inline bool estimated_hit(const DevVolume& v, float alpha, v3 pos, v3 dir, float dis, v3* next, int* taps) {
    float smax = v.max_density * alpha;
    float t = 0.0f, last_non_empty_t = 0.0f, prob = 1.0f;
    float t_step = 2.0f / (float)v.N;
    float pass_rate = expf(-t_step * smax);
    float inv_res = 1.0f / (float)v.N;
    int loop = 0;
    float last_dens = 0.0f;
    BrickCache bc;
    while (loop++ < 10000) {
        if (t > dis) {
            if (last_non_empty_t != 0.0f) { t = last_non_empty_t; break; }
            *next = mk3(0, 0, 0);
            return false;
        }
        const v3 sp = pos + (dir * t);
        const FullTap tp = full_tap(v.N, sp);
        float d = (v.occ && tap_in_empty_brick(v, tp, bc)) ? 0.0f : density_full_at(v, tp);
        ++*taps;
        if (d > 0.0f) {
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
            if (last_dens > 0.0f) {
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
            t += fmaxf(-d * 1.0f, inv_res);
        }
        last_dens = d;
    }
    *next = (dir * t) + pos;
    return true;
}

// ---------------------------------------------------------------------------------------
// Optical-depth march of the tile feature gather: DeterminateEstimatedHit float4 version with
// CheckExitCondition (render.cuh:199-397) and GetSampleT (render.cuh:464-486).  Trapezoid
// optical depth along the ray; ts = (t at survival 0.75, t at survival 0.25, first non-empty t,
// opacity).  Same arithmetic as the oracle's estimated_hit_tau.
// ---------------------------------------------------------------------------------------
// -logf(0.75f), -logf(0.76f), -logf(0.25f), correctly rounded
constexpr float kTauFirst = 0x1.269622p-2f, kTauStepMax = 0x1.1905fap-2f, kTauSecond = 0x1.62e43p+0f;
struct Ts4 {
    float x, y, z, w;
};
__device__ inline float trapezoid_offset(float A, float B, float C, float dt, float S, float lastS, float delta_tau) {
    if (fabsf(A) < 1e-6f) return delta_tau / (dt * (S + lastS) * 0.5f);  // linear fallback
    float D = B * B - 4.0f * A * C;
    D = fmaxf(D, 0.0f);
    const float sqrtD = sqrtf(D);
    const float q = -0.5f * (B + sqrtD);
    const float u1 = q / A;
    const float u2 = C / q;
    return u1 == fminf(fmaxf(u1, 0.0f), 1.0f) ? u1 : u2;
}
__device__ inline bool check_exit(float S, float lastS, float dt, float t, float& tau, float& tau_target, Ts4& ts,
                                  bool& first_hit) {
    const float tau_next = tau + (S + lastS) * 0.5f * dt;
    if (tau_next >= tau_target) {
        const float delta_tau = tau_target - tau;
        const float A = 0.5f * dt * (S - lastS);
        const float B = dt * lastS;
        const float C = -delta_tau;
        float offset = trapezoid_offset(A, B, C, dt, S, lastS, delta_tau);
        offset = fminf(fmaxf(offset, 0.0f), 1.0f);
        const float hit_t = t - dt * (1.0f - offset);
        if (first_hit) {
            ts.x = hit_t;
            tau_target = kTauSecond;
            first_hit = false;
            if (tau_next >= tau_target) {
                const float delta_tau2 = tau_target - tau;
                const float offset2 = trapezoid_offset(A, B, -delta_tau2, dt, S, lastS, delta_tau2);  // not clamped
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
__device__ inline Ts4 estimated_hit_tau(const DevVolume& v, float alpha, v3 pos, v3 dir, float dis) {
    const float smax = v.max_density * alpha;
    float t = 0.0f;
    Ts4 ts = {0.0f, 0.0f, 3.402823466e38f, 0.0f};
    float last_non_empty_t = 0.0f;
    float t_step = 1.0f / (float)v.N;
    if (smax * t_step > kTauStepMax) t_step = kTauStepMax / smax;
    float tau = 0.0f, tau_target = kTauFirst;
    bool first_hit = true;
    int loop = 0;
    float last_dens = 0.0f;
    BrickCache bc;
    while (loop++ < 10000) {
        if (t > dis) {
            if (first_hit) ts.x = last_non_empty_t;
            else ts.y = last_non_empty_t;
            break;
        }
        const FullTap tp = full_tap(v.N, pos + (dir * t));
        const float d = (v.occ && tap_in_empty_brick(v, tp, bc)) ? 0.0f : density_full_at(v, tp);
        if (d > 0.0f) {
            const float S = d * alpha;
            if (check_exit(S, last_dens, t_step, t, tau, tau_target, ts, first_hit)) break;
            ts.z = fminf(t, ts.z);
            t += t_step;
            last_dens = S;
        } else {
            if (last_dens > 0.0f) {
                const float st = t_step + d;  // a negative SDF value shortens the step
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
    if (ts.z == 3.402823466e38f) ts.z = 0.0f;
    const float survival = expf(-tau);
    ts.w = fminf(fmaxf(1.0f - survival, 0.0f) / 0.75f, 1.0f);
    return ts;
}
__device__ inline Ts4 get_sample_t(const DevVolume& v, v3 ori, v3 dir, float alpha) {
    const float dis = ray_box_offset(ori, dir);
    if (dis < 0.0f) return Ts4{0.0f, 0.0f, 0.0f, 0.0f};
    const v3 sp = ori + dir * dis;
    const float max_dis = ray_box_distance(sp, dir);
    Ts4 ts = estimated_hit_tau(v, alpha, sp, dir, max_dis);
    if (ts.x > 0.0f) {
        ts.x += dis;
        ts.z += dis;
    }
    if (ts.y > 0.0f) ts.y += dis;
    return ts;
}

// XCD-aware block numbering.  Workgroups are dispatched round-robin over the 8 XCDs of an
// MI355X (block b runs on XCD b % 8) and each XCD has its own L2: renumbering blocks so that
// every XCD owns one contiguous range of work items keeps neighbouring pixels (which read
// the same texels) on the same L2.  Grids not divisible by 8 keep the plain order.
constexpr int kXcds = 8;
__device__ inline int xcd_swizzle(int b, int nb) {
    if (nb % kXcds) return b;
    return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// Per-pixel state after ray generation, march and basis (volume.cu:198-216,
// render.cuh:808-823).  Pixel jitter and basis draws use Philox streams.
struct ViewDev {
    v3 ori, up, right, light_dir, light_color, scatter_rate;
    float alpha, g, tan_half_fov, jitter;
    int width, height;
    uint32_t seed, frame, view_id;
    int volume_index;
};
struct PixelState {
    v3 pos, X, Y, Z, LX, L;
    float dist;
    int active;
    int taps;
};

__device__ inline v3 pixel_ray(const ViewDev& vw, int row, int col, uint32_t pix) {
    float xu, xv;
    if (vw.jitter < 0.0f) {
        PhiloxStream rs(vw.seed, vw.frame, pix, RNG_JITTER, vw.view_id);
        xu = rs.next();
        xv = rs.next();
    } else {
        xu = xv = vw.jitter;
    }
    float u = 1.0f - ((float)col + xu) / (float)vw.width;
    float vv = 1.0f - ((float)row + xv) / (float)vw.height;
    float sy = vw.tan_half_fov;
    float sx = vw.tan_half_fov * ((float)vw.width / (float)vw.height);
    v3 fwd = normalize3(-vw.ori);
    v3 dir = fwd + (vw.right * ((u * 2.0f - 1.0f) * sx)) + (vw.up * ((vv * 2.0f - 1.0f) * sy));
    return normalize3(dir);  // volume.cu:216
}

// GetMatrixFromNormal(seed, v1, 0.1) (render.cuh:694-708); at most 64 attempts.
__device__ inline void basis_from_normal(v3 v1, const ViewDev& vw, uint32_t pix, v3* X, v3* Y, v3* Z) {
    v1 = normalize3(v1);
    PhiloxStream rs(vw.seed, vw.frame, pix, RNG_BASIS_MAIN, vw.view_id);
    v3 v2 = mk3(0, 1, 0), v3_ = mk3(0, 0, 1);
    for (int att = 0; att < 64; ++att) {
        float e0 = rs.next() * 0.1f;
        float e1 = rs.next() * 0.1f;
        v3 r = uniform_sample_sphere(e0, e1);
        // the reference's double compare |d| < 1e-5: float(1e-5) is the largest float below 1e-5
        if (fabsf(dot3(r, v1)) <= 0.00001f) continue;
        v2 = normalize3(cross3(v1, r));
        v3_ = cross3(v2, v1);
        break;
    }
    *X = v1;
    *Y = v2;
    *Z = v3_;
}

__device__ inline PixelState pixel_state(const DevVolume& vol, const ViewDev& vw, int row, int col) {
    PixelState st;
    uint32_t pix = (uint32_t)(row * vw.width + col);
    v3 dir = pixel_ray(vw, row, col, pix);
    dir = normalize3(dir);                  // render.cuh:811
    v3 light = normalize3(vw.light_dir);    // render.cuh:812
    // GetSamplePoint (render.cuh:436-462)
    v3 mdir = normalize3(dir);
    st.active = 0;
    st.taps = 0;
    st.pos = mk3(0, 0, 0);
    float dis = ray_box_offset(vw.ori, mdir);
    if (dis >= 0.0f) {
        v3 sp = vw.ori + mdir * dis;
        float maxd = ray_box_distance(sp, mdir);
        v3 p0;
        if (estimated_hit(vol, vw.alpha, sp, mdir, maxd, &p0, &st.taps)) {
            st.active = 1;
            st.pos = p0;
        }
    }
    basis_from_normal(dir, vw, pix, &st.X, &st.Y, &st.Z);
    st.LX = normalize3(light);
    st.L = light;
    st.dist = length3(vw.ori - st.pos);
    return st;
}

// One stencil point of SetupInputData (radiancePredict.cu:433-471): returns
// (X_Val, X_Val_Hg, X_Val_Sub) of point i.
__device__ inline v3 gather_point(const DevVolume& vol, const StencilEntry& s, int i, const PixelState& st, float g,
                                  float alpha) {
    v3 p;
    if (s.type >= 4) p = st.pos + ((st.X * s.dx + st.Y * s.dy + st.Z * s.dz) * s.offset);
    else p = st.pos + st.L * s.offset;
    float xv = logf(1.0f + alpha * mip_density(vol, s.mip, p) / 64.0f);
    float xh;
    if (s.local == 0) {
        xh = logf(henyey_greenstein(dot3(st.X, st.L), g) + 1.0f);
    } else {
        v3 ms = normalize3(p - st.pos);
        float c1 = dot3(st.X, ms), c2 = dot3(st.LX, ms);
        float u = c1 * 0.5f + 0.5f, u2 = c2 * 0.5f + 0.5f;
        float h0 = sample_lut_exact(vol.lut, u, s.lut_v);
        float h1 = sample_lut_exact(vol.lut, u2, s.lut_v);
        xh = logf(h0 * h1 + 1.0f);
    }
    float xs = shadow_term_tr(vol, p, st.LX, s.mip - 1 > 0 ? s.mip - 1 : 0);
    return mk3(xv, xh, xs);
}

__device__ inline float gamma_of(const PixelState& st) {  // radiancePredict.cu:482-484
    float d = dot3(st.X, st.LX);
    d = fminf(1.0f, fmaxf(-1.0f, d));
    return acosf(d);
}

}  // namespace mrp
