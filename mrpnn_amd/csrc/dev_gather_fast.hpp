// Fast feature gather of the fused MFMA kernels.
//
// Same 192 sample points, filters and feature formulas as SetupInputData
// (core/radiancePredict.cu:427-485, exact form in dev_common.hpp gather_point), organised for
// throughput:
//   * density mips 0..7 and TR mips 0..6 are read from zero-padded copies ((res+3)^3 per
//     level, data at [1,res] per axis): the continuous coordinate is clamped into the padded
//     range, so every tap is in bounds and no per-tap border test is needed; the clamp also
//     reproduces the border (density) and clamp (TR) addressing exactly;
//   * the padded levels are stored as (y, z) quads (texel (x, y, z) holds its own value and those
//     at z+1, y+1 and y+1, z+1), so a trilinear blend is two 16-byte buffer loads (slices x and
//     x+1) at 32-bit offsets from a uniform buffer resource, instead of four 8-byte (z, z+1)
//     pairs: half the load instructions, whose issue the kernel is short of;
//   * the 144 points of stencil chunks 16..159 share one density mip per 16-point chunk, so
//     their level geometry is a compile-time constant (gather_chunk): no per-point table reads
//     of it, the texel index is two 24-bit multiply-adds by constants;
//   * the per-point constants (direction * offset, 1/|direction|, LUT v, mip geometry) come
//     from a 192-entry table; for sphere points the first LUT lookup depends only on the
//     point (the basis is orthonormal, dot(XMain, direction) = dx/|d|), so it is tabulated
//     once per LUT;
//   * lerps are a + w*(b - a) with one FMA; ln(1 + x) of the non-negative feature terms uses
//     the hardware log2.
// Results differ from the exact path by a few ulps of the features.
#pragma once
#include "dev_common.hpp"

namespace mrp {

// Padded level geometry (shared by the host allocation and the kernels).
MRP_HD constexpr int pad_dim(int res) { return res + 3; }
MRP_HD constexpr long long pmip_offset(int m) {  // density mips 0..7, base 256
    long long o = 0;
    for (int i = 0; i < m; ++i) o += (long long)pad_dim(256 >> i) * pad_dim(256 >> i) * pad_dim(256 >> i);
    return o;
}
MRP_HD constexpr long long ptr_offset(int m) {  // TR mips 0..6, base 128
    long long o = 0;
    for (int i = 0; i < m; ++i) o += (long long)pad_dim(128 >> i) * pad_dim(128 >> i) * pad_dim(128 >> i);
    return o;
}
constexpr long long kPadMipFloats = pmip_offset(8);
constexpr long long kPadTrFloats = ptr_offset(7);
// (y, z) difference quads of the padded density mips 0..7 (the fast gather's and the TR bake's
// sources): texel (x, y, z) of a padded level holds (p(x,y,z), p(x,y,z+1) - p(x,y,z), p(x,y+1,z),
// p(x,y+1,z+1) - p(x,y+1,z)), so a trilinear blend is two 16-byte loads (slices x and x + 1).
// Offsets in quads; a level's quads sit at the texel offsets of the padded mips.
MRP_HD constexpr long long qmip_offset(int m) { return pmip_offset(m); }
constexpr long long kQuadTexels = qmip_offset(8);
// (y, z) quads of the zero-padded, edge-replicated TR mips 0..6 (the fast gather's TR blends),
// written by the bake; offsets in quads, at the texel offsets of the padded levels.
MRP_HD constexpr long long qtr_offset(int m) { return ptr_offset(m); }
constexpr long long kQuadTrTexels = qtr_offset(7);
constexpr long long kPadTrAlloc = 4 * kQuadTrTexels;  // floats of a TR set's quad allocation

// One stencil point of the fast gather (80 bytes).
struct FastStencil {
    float ddx, ddy, ddz;  // table direction * offset (sphere points)
    float cscale;         // 1 / (|table direction| * offset) (sphere points)
    float off;            // offset (light points)
    float hg0;            // LUT(dx/|d| * 0.5 + 0.5, v) (sphere points)
    float dres, tres;     // density / TR mip resolution
    unsigned dRu, tRu;    // padded dims
    unsigned dbase, tbase;  // padded level offsets (bytes)
    unsigned dR4, dRR4;     // 4 * R, 4 * R * R (byte strides of the density level)
    unsigned tR4, tRR4;     // same for the TR level
    unsigned lbase;         // byte offset of this point's LUT row (LUT rows at v = lut_v)
    int kind;               // 0 centre, 1 sphere, 2 light
    float lut_v;            // LUT v coordinate
    int pad0;
};
static_assert(sizeof(FastStencil) == 80, "FastStencil must be 80 bytes");

// LUT rows: for every stencil point the 512-texel row of the HG LUT interpolated at the
// point's v (the bilinear lookup is separable, so a lookup becomes one linear interpolation);
// entry 512 repeats 511 so a pair load at the clamped edge stays in the row.
constexpr int kLutRow = 516;

// Buffer resource over a device array: loads are bounds-checked by the hardware (reads
// past num_records return 0) and take a 32-bit byte offset, so a trilinear tap pair costs one
// 64-bit load and no 64-bit address arithmetic.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ inline BufRsrc make_rsrc(const float* p, long long n_floats) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n_floats * 4), 0x00020000);
}
__device__ inline f32x2_t ld_pair(BufRsrc r, int byte_off) {
    return __builtin_bit_cast(f32x2_t, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}

// Per-tile (uniform) inputs.
struct FastView {
    BufRsrc rows;    // LUT rows
    BufRsrc qm, qt;  // (y, z) quads of the padded density mips 0..7 and TR mips 0..6
    v3 L, LX, invLX;  // light direction as the gather uses it; LX = normalize(L); inv3(LX)
    v3 a2f;           // 0.4999 * invLX (the box half-width over the light direction)
    float alpha64;    // alpha / 64
};
// Per-pixel inputs.
struct FastPixel {
    v3 pos, X, Y, Z;
    float ax, ay, az;  // dot(LX, X), dot(LX, Y), dot(LX, Z)
    float uc1, uc2;    // LUT u of light points: dot(X,L)*.5+.5, dot(LX,L)*.5+.5
    float xh_centre;   // X_Val_Hg of point 0: log(HG(dot(X, L), g) + 1)
};

__device__ inline float flerp(float a, float b, float w) { return fmaf(w, b - a, a); }
__device__ inline float clampf(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
// ln(x) for x >= 1 (no denormal inputs): hardware log2 times ln 2.
__device__ inline float log_ge1(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }

// Corner and weights of a trilinear blend on a padded level of R^3 texels at the non-negative
// continuous coordinate (cx, cy, cz) (already inside the padding): corner = floor (a truncating
// conversion, the coordinate being non-negative), weights = fract; o4 = 4 * ((x*R + y)*R + z)
// (bytes, without the level base), formed with 24-bit integer multiply-adds: the padded 256^3
// level has 259^3 > 2^24 texels, beyond the integers a float holds exactly.  Levels of equal
// resolution share it.
struct TapGeo {
    unsigned o4;
    float wx, wy, wz;
};
__device__ inline TapGeo tap_geo(unsigned R, float cx, float cy, float cz) {
    TapGeo g;
    g.wx = __builtin_amdgcn_fractf(cx);
    g.wy = __builtin_amdgcn_fractf(cy);
    g.wz = __builtin_amdgcn_fractf(cz);
    // through int: defined for the negative coordinates of lanes whose result is discarded
    // (their index is then any 32-bit value; such lanes only occur where the caller clamps)
    const unsigned x = (unsigned)(int)cx, y = (unsigned)(int)cy, z = (unsigned)(int)cz;
    g.o4 = mad_u24_v(mad_u24_v(x, R, y), R, z) * 4u;
    return g;
}
// The same with a compile-time R (scalar operand).
template <unsigned R>
__device__ inline TapGeo tap_geo_k(float cx, float cy, float cz) {
    TapGeo g;
    g.wx = __builtin_amdgcn_fractf(cx);
    g.wy = __builtin_amdgcn_fractf(cy);
    g.wz = __builtin_amdgcn_fractf(cz);
    const unsigned x = (unsigned)(int)cx, y = (unsigned)(int)cy, z = (unsigned)(int)cz;
    g.o4 = mad_u24_s(mad_u24_s(x, R, y), R, z) * 4u;
    return g;
}
// Trilinear blend on a level at byte offset BASE with strides R4 = 4R, RR4 = 4R^2, all
// compile-time.  (Addressing the four tap pairs through the instruction's scalar offset instead
// of three vector adds was measured 5 % slower on the network kernel.)
template <unsigned BASE, unsigned R4, unsigned RR4>
__device__ inline float tri_k(BufRsrc r, const TapGeo& g) {
    const int o = (int)(g.o4 + BASE);
    const f32x2_t c00 = ld_pair(r, o), c01 = ld_pair(r, o + (int)R4);
    const f32x2_t c10 = ld_pair(r, o + (int)RR4), c11 = ld_pair(r, o + (int)(RR4 + R4));
    return flerp(flerp(flerp(c00.x, c00.y, g.wz), flerp(c01.x, c01.y, g.wz), g.wy),
                 flerp(flerp(c10.x, c10.y, g.wz), flerp(c11.x, c11.y, g.wz), g.wy), g.wx);
}
// The same blend from (y, z) quads (16 bytes per padded texel) at byte offset BASE, RR16 = 16 R^2:
// two 16-byte loads (slices x and x + 1) instead of four 8-byte pairs; same taps, blend and
// rounding.  Plain quads hold (p(y,z), p(y,z+1), p(y+1,z), p(y+1,z+1)); difference quads (DIFF,
// the density quads) hold (p(y,z), p(y,z+1) - p(y,z), p(y+1,z), p(y+1,z+1) - p(y+1,z)), the
// differences the z lerps a + w (b - a) form, computed once when the quads are built: those four
// lerps are then one FMA each, with the same bits.
template <bool DIFF>
__device__ inline float quad_blend(const f32x4_t& q0, const f32x4_t& q1, const TapGeo& g) {
    if (DIFF)
        return flerp(flerp(fmaf(g.wz, q0.y, q0.x), fmaf(g.wz, q0.w, q0.z), g.wy),
                     flerp(fmaf(g.wz, q1.y, q1.x), fmaf(g.wz, q1.w, q1.z), g.wy), g.wx);
    return flerp(flerp(flerp(q0.x, q0.y, g.wz), flerp(q0.z, q0.w, g.wz), g.wy),
                 flerp(flerp(q1.x, q1.y, g.wz), flerp(q1.z, q1.w, g.wz), g.wy), g.wx);
}
template <unsigned BASE, unsigned RR16, bool DIFF>
__device__ inline float tri_q(BufRsrc r, const TapGeo& g) {
    const int o = (int)(g.o4 * 4u + BASE);
    const f32x4_t q0 = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0));
    const f32x4_t q1 = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, o + (int)RR16, 0, 0));
    return quad_blend<DIFF>(q0, q1, g);
}
// The same with the level's geometry at run time: qbase = 16 x the level's texel offset, RR16 =
// 16 R^2 (the byte offsets and strides of the pair layout, times four).
template <bool DIFF>
__device__ inline float tri_q_rt(BufRsrc r, unsigned qbase, unsigned RR16, const TapGeo& g) {
    const int o = (int)(g.o4 * 4u + qbase);
    const f32x4_t q0 = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0));
    const f32x4_t q1 = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, o + (int)RR16, 0, 0));
    return quad_blend<DIFF>(q0, q1, g);
}
__device__ inline float tri_geo(BufRsrc r, unsigned base, unsigned R4, unsigned RR4, const TapGeo& g) {
    const int o = (int)(g.o4 + base);
    const f32x2_t c00 = ld_pair(r, o), c01 = ld_pair(r, o + (int)R4);
    const f32x2_t c10 = ld_pair(r, o + (int)RR4), c11 = ld_pair(r, o + (int)(RR4 + R4));
    return flerp(flerp(flerp(c00.x, c00.y, g.wz), flerp(c01.x, c01.y, g.wz), g.wy),
                 flerp(flerp(c10.x, c10.y, g.wz), flerp(c11.x, c11.y, g.wz), g.wy), g.wx);
}
__device__ inline float tri_taps(BufRsrc r, unsigned base, unsigned R, unsigned R4, unsigned RR4, float cx, float cy,
                                 float cz) {
    return tri_geo(r, base, R4, RR4, tap_geo(R, cx, cy, cz));
}
// Padded continuous coordinate of p on a level of resolution res: data coordinate
// (p + 0.5)*res - 0.5, plus one for the padding.
__device__ inline v3 padded_coord(v3 p, float res) {
    const float k = fmaf(0.5f, res, 0.5f);
    return mk3(fmaf(p.x, res, k), fmaf(p.y, res, k), fmaf(p.z, res, k));
}
// Zero-border trilinear (density): padded coordinate clamped to [0, res+1] (both taps
// outside -> exact 0).
__device__ inline float tri_border_c(BufRsrc r, const FastStencil& s, v3 c) {
    const float hi = s.dres + 1.0f;
    return tri_taps(r, s.dbase, s.dRu, s.dR4, s.dRR4, clampf(c.x, 0.0f, hi), clampf(c.y, 0.0f, hi),
                    clampf(c.z, 0.0f, hi));
}
// Clamp-addressing trilinear (TR mips): data coordinate clamped to [0, res-1], i.e. padded
// coordinate to [1, res].
__device__ inline float tri_clamp_c(BufRsrc r, const FastStencil& s, v3 c) {
    const float hi = s.tres;
    return tri_taps(r, s.tbase, s.tRu, s.tR4, s.tRR4, clampf(c.x, 1.0f, hi), clampf(c.y, 1.0f, hi),
                    clampf(c.z, 1.0f, hi));
}
// Point-level LUT lookup: linear interpolation along u of the point's LUT row.
__device__ inline float lut_row(BufRsrc rows, unsigned lbase, float u) {
    const float px = clampf(fmaf(u, 512.0f, -0.5f), 0.0f, 511.0f);
    const f32x2_t a = ld_pair(rows, (int)((unsigned)px * 4u + lbase));
    return flerp(a.x, a.y, __builtin_amdgcn_fractf(px));
}
// HenyeyGreenstein (vector.cu:210-214) with t^1.5 formed as t * sqrt(t) (two correctly
// rounded steps instead of a powf call).
__device__ inline float hg_fast(float cs, float g) {
    const float g2 = g * g;
    const float t = 1.0f + g2 - 2.0f * g * cs;
    return ((1.0f - g2) / (t * sqrtf(t))) * 0.25f;
}

// Zero-border trilinear on padded level data of resolution res at byte offset base.
__device__ inline float tri_border_lvl(BufRsrc r, unsigned base, float res, v3 p) {
    const unsigned Ru = (unsigned)(res + 3.0f);
    const v3 c = padded_coord(p, res);
    const float hi = res + 1.0f;
    return tri_taps(r, base, Ru, Ru * 4u, Ru * Ru * 4u, clampf(c.x, 0.0f, hi), clampf(c.y, 0.0f, hi),
                    clampf(c.z, 0.0f, hi));
}
// 512x512 LUT, bilinear, clamp (used to build the rows).
__device__ inline float lut_bilinear(const float* __restrict__ lut, float u, float v) {
    const float px = clampf(fmaf(u, 512.0f, -0.5f), 0.0f, 511.0f);
    const float py = clampf(fmaf(v, 512.0f, -0.5f), 0.0f, 511.0f);
    const float wx = __builtin_amdgcn_fractf(px), wy = __builtin_amdgcn_fractf(py);
    const unsigned x0 = (unsigned)px, y0 = (unsigned)py;
    const unsigned x1 = min(x0 + 1u, 511u), y1 = min(y0 + 1u, 511u);
    return flerp(flerp(lut[y0 * 512u + x0], lut[y0 * 512u + x1], wx), flerp(lut[y1 * 512u + x0], lut[y1 * 512u + x1], wx),
                 wy);
}

// ShadowTerm_TRTex(...).x (renderResources.cu:84-102) on the padded TR level (replicated
// edges).  g_same: the density blend's corner and weights at p when the TR level has the
// density level's resolution (stencil mips >= 1); for p inside the box that coordinate lies in
// [0.5, res + 0.5], where the replicated edges give the clamp-addressed value, so the TR
// blend reuses it.  Other points take their own clamped coordinate (the box entry along the
// light for points outside); one blend serves every lane.
// Outside the box [-0.5, 0.5]^3 (the test of renderResources.cu:86-87): one max3 of |p|.
__device__ inline bool outside_box(v3 p) { return fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))) > 0.5f; }
__device__ inline float shadow_fast(const FastView& fv, const FastStencil& s, v3 p, const TapGeo& g_same,
                                    bool same, bool outside) {
    TapGeo g = g_same;
    bool miss = false;
    // wave-uniform test first: when every lane of a `same` point is inside the box (the common
    // case) the own-coordinate path is skipped as a whole
    if (!same || __builtin_amdgcn_ballot_w64(outside) != 0)
    if (outside || !same) {
        v3 q = p;
        if (outside) {
            // RayBoxOffset(p, LX) (vector.cu:46-63) with inv(LX) precomputed
            const v3 to_axil = (-p) * fv.invLX;
            const v3 d0 = to_axil + fv.a2f, d1 = to_axil - fv.a2f;
            const v3 tmin = vmin3(d0, d1), tmax = vmax3(d0, d1);
            const float tmi = fmaxf(tmin.x, fmaxf(tmin.y, tmin.z));
            const float tma = fminf(tmax.x, fminf(tmax.y, tmax.z));
            miss = !(tma >= tmi);
            const float off = fmaxf(tmi, 0.0f);
            q = mk3(fmaf(fv.LX.x, off, p.x), fmaf(fv.LX.y, off, p.y), fmaf(fv.LX.z, off, p.z));
        }
        const v3 c = padded_coord(q, s.tres);
        g = tap_geo(s.tRu, clampf(c.x, 1.0f, s.tres), clampf(c.y, 1.0f, s.tres), clampf(c.z, 1.0f, s.tres));
    }
    // quads: the pair layout's byte offsets and strides times four
    const float T = tri_q_rt<false>(fv.qt, s.tbase * 4u, s.tRR4 * 4u, g);
    return miss ? 1.0f : T;
}

// (X_Val, X_Val_Hg, X_Val_Sub) of one stencil point.  LIGHT: points 160..191 (along the
// light); centre: point 0; same: the point's density and TR levels have equal resolution.
// LIGHT is a template argument and same a constant after inlining at every call site.
template <bool LIGHT>
__device__ inline v3 gather_fast(const FastView& fv, const FastStencil& s, const FastPixel& px, bool centre,
                                 bool same) {
    v3 p;
    float xh;
    if (LIGHT) {  // MsDir = normalize(L * off) = L
        p = mk3(fmaf(fv.L.x, s.off, px.pos.x), fmaf(fv.L.y, s.off, px.pos.y), fmaf(fv.L.z, s.off, px.pos.z));
        const float h0 = lut_row(fv.rows, s.lbase, px.uc1), h1 = lut_row(fv.rows, s.lbase, px.uc2);
        xh = log_ge1(fmaf(h0, h1, 1.0f));
    } else {
        p.x = fmaf(px.X.x, s.ddx, fmaf(px.Y.x, s.ddy, fmaf(px.Z.x, s.ddz, px.pos.x)));
        p.y = fmaf(px.X.y, s.ddx, fmaf(px.Y.y, s.ddy, fmaf(px.Z.y, s.ddz, px.pos.y)));
        p.z = fmaf(px.X.z, s.ddx, fmaf(px.Y.z, s.ddy, fmaf(px.Z.z, s.ddz, px.pos.z)));
        // cos between LX and the point direction: (ax*dx + ay*dy + az*dz)/|d|, ddx = dx*off
        const float c2 = fmaf(px.ax, s.ddx, fmaf(px.ay, s.ddy, px.az * s.ddz)) * s.cscale;
        const float h1 = lut_row(fv.rows, s.lbase, fmaf(c2, 0.5f, 0.5f));
        xh = centre ? px.xh_centre : log_ge1(fmaf(s.hg0, h1, 1.0f));
    }
    const v3 c = padded_coord(p, s.dres);
    const bool outside = outside_box(p);
    // inside the box the padded coordinate is in [0.5, res + 0.5]: no clamp needed; lanes
    // outside clamp to the zero border [0, res + 1] (skipped when no lane of the wave is outside)
    TapGeo g = tap_geo(s.dRu, c.x, c.y, c.z);
    if (__builtin_amdgcn_ballot_w64(outside) != 0 && outside) {
        const float hi = s.dres + 1.0f;
        g = tap_geo(s.dRu, clampf(c.x, 0.0f, hi), clampf(c.y, 0.0f, hi), clampf(c.z, 0.0f, hi));
    }
    const float rho = tri_q_rt<true>(fv.qm, s.dbase * 4u, s.dRR4 * 4u, g);
    const float xv = log_ge1(fmaf(fv.alpha64, rho, 1.0f));
    const float xs = shadow_fast(fv, s, p, g, same, outside);
    return mk3(xv, xh, xs);
}

// Light-ray entry of a point outside the box (RayBoxOffset, vector.cu:46-63, with inv(LX) and
// 0.4999 * inv(LX) precomputed): the entry point q, miss = the ray passes the box.
__device__ inline v3 light_entry(const FastView& fv, v3 p, bool& miss) {
    const v3 to_axil = (-p) * fv.invLX;
    const v3 d0 = to_axil + fv.a2f, d1 = to_axil - fv.a2f;
    const v3 tmin = vmin3(d0, d1), tmax = vmax3(d0, d1);
    const float tmi = fmaxf(tmin.x, fmaxf(tmin.y, tmin.z));
    const float tma = fminf(tmax.x, fminf(tmax.y, tmax.z));
    miss = !(tma >= tmi);
    const float off = fmaxf(tmi, 0.0f);
    return mk3(fmaf(fv.LX.x, off, p.x), fmaf(fv.LX.y, off, p.y), fmaf(fv.LX.z, off, p.z));
}

// (X_Val, X_Val_Hg, X_Val_Sub) of a sphere point of a chunk whose density mip M (2..7) is a
// compile-time constant; TR level M - 1 has the same resolution.  The same arithmetic as
// gather_fast<false>(..., centre = false, same = true); only the addressing differs.
template <int M>
__device__ inline v3 gather_chunk(const FastView& fv, const FastStencil& s, const FastPixel& px) {
    static_assert(M >= 2 && M <= 7, "uniform chunks use density mips 2..7");
    constexpr unsigned res = 256u >> M, R = res + 3u;
    // byte offsets of the level in the density and TR quads; 16 R^2 bytes per x slice
    constexpr unsigned QDB = (unsigned)qmip_offset(M) * 16u, QTB = (unsigned)qtr_offset(M - 1) * 16u;
    constexpr unsigned RR16 = 16u * R * R;
    static_assert(res == (128u >> (M - 1)), "TR level M - 1 has the density level's resolution");
    constexpr float fres = (float)res;
    v3 p;
    p.x = fmaf(px.X.x, s.ddx, fmaf(px.Y.x, s.ddy, fmaf(px.Z.x, s.ddz, px.pos.x)));
    p.y = fmaf(px.X.y, s.ddx, fmaf(px.Y.y, s.ddy, fmaf(px.Z.y, s.ddz, px.pos.y)));
    p.z = fmaf(px.X.z, s.ddx, fmaf(px.Y.z, s.ddy, fmaf(px.Z.z, s.ddz, px.pos.z)));
    const float c2 = fmaf(px.ax, s.ddx, fmaf(px.ay, s.ddy, px.az * s.ddz)) * s.cscale;
    const float h1 = lut_row(fv.rows, s.lbase, fmaf(c2, 0.5f, 0.5f));
    const float xh = log_ge1(fmaf(s.hg0, h1, 1.0f));
    v3 c = padded_coord(p, fres);
    const bool outside = outside_box(p);
    const bool any_out = __builtin_amdgcn_ballot_w64(outside) != 0;
    // zero border: clamp to [0, res + 1] (inside the box the coordinate is in [0.5, res + 0.5],
    // which the clamp leaves alone; the rings of mips >= 5 reach past the box in most
    // wavefronts, so they clamp unconditionally)
    if (M >= 5 || (any_out && outside))
        c = mk3(clampf(c.x, 0.0f, fres + 1.0f), clampf(c.y, 0.0f, fres + 1.0f), clampf(c.z, 0.0f, fres + 1.0f));
    const TapGeo g = tap_geo_k<R>(c.x, c.y, c.z);
    const float rho = tri_q<QDB, RR16, true>(fv.qm, g);
    const float xv = log_ge1(fmaf(fv.alpha64, rho, 1.0f));
    // ShadowTerm_TRTex: inside the box the density blend's corner and weights (replicated TR
    // edges); outside, the light ray's entry point, clamp addressing
    TapGeo gt = g;
    bool miss = false;
    if (any_out && outside) {
        const v3 q = light_entry(fv, p, miss);
        const v3 ct = padded_coord(q, fres);
        gt = tap_geo_k<R>(clampf(ct.x, 1.0f, fres), clampf(ct.y, 1.0f, fres), clampf(ct.z, 1.0f, fres));
    }
    const float T = tri_q<QTB, RR16, false>(fv.qt, gt);
    return mk3(xv, xh, miss ? 1.0f : T);
}

}  // namespace mrp
