// Fast feature gather of the fused MFMA kernels.
//
// Same 192 sample points, filters and feature formulas as SetupInputData
// (core/radiancePredict.cu:427-485, exact form in dev_common.hpp gather_point), organised for
// throughput:
//   * density mips 0..7 and TR mips 0..6 are read from zero-padded copies ((res+3)^3 per
//     level, data at [1,res] per axis): the continuous coordinate is clamped into the padded
//     range, so every tap is in bounds and no per-tap border test is needed; the clamp also
//     reproduces the border (density) and clamp (TR) addressing exactly;
//   * every tap is one uniform base pointer plus a 32-bit offset;
//   * the per-point constants (direction * offset, 1/|direction|, LUT v, mip geometry) come
//     from a 192-entry table; for sphere points the first LUT lookup depends only on the
//     point (the basis is orthonormal, dot(XMain, direction) = dx/|d|), so it is tabulated
//     once per LUT;
//   * lerps are a + w*(b - a) with one FMA.
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

// One stencil point of the fast gather (64 bytes).
struct FastStencil {
    float ddx, ddy, ddz;  // table direction * offset (sphere points)
    float cscale;         // 1 / (|table direction| * offset) (sphere points)
    float off;            // offset (light points)
    float lut_v;          // LUT v coordinate
    float hg0;            // LUT(dx/|d| * 0.5 + 0.5, v) (sphere points)
    float dres, tres;     // density / TR mip resolution
    unsigned dbase, tbase;  // padded level offsets (floats)
    int dR, tR;             // padded dims
    int kind;               // 0 centre, 1 sphere, 2 light
    int pad0, pad1;
};
static_assert(sizeof(FastStencil) == 64, "FastStencil must be 64 bytes");

// Per-tile (uniform) inputs.
struct FastView {
    const float* __restrict__ pmips;
    const float* __restrict__ ptr;
    const float* __restrict__ lut;
    v3 L, LX, invLX;  // light direction as the gather uses it; LX = normalize(L); inv3(LX)
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

// Zero-border trilinear on a padded level: uvw = p + 0.5, data coordinate uvw*res - 0.5,
// padded coordinate +1, clamped to [0, res+1] (both taps outside -> exact 0).
__device__ inline float tri_border(const float* __restrict__ base, unsigned lvl, int R, float res, v3 p) {
    const float k = fmaf(0.5f, res, 0.5f), hi = res + 1.0f;
    const float cx = fminf(fmaxf(fmaf(p.x, res, k), 0.0f), hi);
    const float cy = fminf(fmaxf(fmaf(p.y, res, k), 0.0f), hi);
    const float cz = fminf(fmaxf(fmaf(p.z, res, k), 0.0f), hi);
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const float wx = cx - fx, wy = cy - fy, wz = cz - fz;
    const unsigned R2 = (unsigned)(R * R);
    const unsigned i = lvl + ((unsigned)fx * (unsigned)R + (unsigned)fy) * (unsigned)R + (unsigned)fz;
    const float* q = base;
    const float c000 = q[i], c001 = q[i + 1], c010 = q[i + R], c011 = q[i + R + 1];
    const float c100 = q[i + R2], c101 = q[i + R2 + 1], c110 = q[i + R2 + R], c111 = q[i + R2 + R + 1];
    return flerp(flerp(flerp(c000, c001, wz), flerp(c010, c011, wz), wy),
                 flerp(flerp(c100, c101, wz), flerp(c110, c111, wz), wy), wx);
}
// Clamp-addressing trilinear (TR mips): data coordinate clamped to [0, res-1].
__device__ inline float tri_clamp(const float* __restrict__ base, unsigned lvl, int R, float res, v3 p) {
    const float k = fmaf(0.5f, res, -0.5f), hi = res - 1.0f;
    const float cx = fminf(fmaxf(fmaf(p.x, res, k), 0.0f), hi) + 1.0f;
    const float cy = fminf(fmaxf(fmaf(p.y, res, k), 0.0f), hi) + 1.0f;
    const float cz = fminf(fmaxf(fmaf(p.z, res, k), 0.0f), hi) + 1.0f;
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const float wx = cx - fx, wy = cy - fy, wz = cz - fz;
    const unsigned R2 = (unsigned)(R * R);
    const unsigned i = lvl + ((unsigned)fx * (unsigned)R + (unsigned)fy) * (unsigned)R + (unsigned)fz;
    const float* q = base;
    const float c000 = q[i], c001 = q[i + 1], c010 = q[i + R], c011 = q[i + R + 1];
    const float c100 = q[i + R2], c101 = q[i + R2 + 1], c110 = q[i + R2 + R], c111 = q[i + R2 + R + 1];
    return flerp(flerp(flerp(c000, c001, wz), flerp(c010, c011, wz), wy),
                 flerp(flerp(c100, c101, wz), flerp(c110, c111, wz), wy), wx);
}
// 512x512 LUT, bilinear, clamp.
__device__ inline float lut_fast(const float* __restrict__ lut, float u, float v) {
    const float px = fminf(fmaxf(fmaf(u, 512.0f, -0.5f), 0.0f), 511.0f);
    const float py = fminf(fmaxf(fmaf(v, 512.0f, -0.5f), 0.0f), 511.0f);
    const float fx = floorf(px), fy = floorf(py);
    const float wx = px - fx, wy = py - fy;
    const unsigned x0 = (unsigned)fx, y0 = (unsigned)fy;
    const unsigned x1 = min(x0 + 1u, 511u), y1 = min(y0 + 1u, 511u);
    const float a = flerp(lut[y0 * 512u + x0], lut[y0 * 512u + x1], wx);
    const float b = flerp(lut[y1 * 512u + x0], lut[y1 * 512u + x1], wx);
    return flerp(a, b, wy);
}

// ShadowTerm_TRTex(...).x (renderResources.cu:84-102) on the padded TR level.
__device__ inline float shadow_fast(const FastView& fv, const FastStencil& s, v3 p) {
    if (p.x < -0.5f || p.y < -0.5f || p.z < -0.5f || p.x > 0.5f || p.y > 0.5f || p.z > 0.5f) {
        // RayBoxOffset(p, LX) (vector.cu:46-63) with inv(LX) precomputed
        const v3 to_axil = (-p) * fv.invLX;
        const v3 a2f = mk3(0.4999f, 0.4999f, 0.4999f) * fv.invLX;
        const v3 d0 = to_axil + a2f, d1 = to_axil - a2f;
        const v3 tmin = vmin3(d0, d1), tmax = vmax3(d0, d1);
        const float tmi = fmaxf(tmin.x, fmaxf(tmin.y, tmin.z));
        const float tma = fminf(tmax.x, fminf(tmax.y, tmax.z));
        if (!(tma >= tmi)) return 1.0f;
        const float off = fmaxf(tmi, 0.0f);
        p = mk3(fmaf(fv.LX.x, off, p.x), fmaf(fv.LX.y, off, p.y), fmaf(fv.LX.z, off, p.z));
    }
    return tri_clamp(fv.ptr, s.tbase, s.tR, s.tres, p);
}

// (X_Val, X_Val_Hg, X_Val_Sub) of one stencil point.  LIGHT: points 160..191 (along the
// light); centre: point 0.  LIGHT is a compile-time constant at every call site.
template <bool LIGHT>
__device__ inline v3 gather_fast(const FastView& fv, const FastStencil& s, const FastPixel& px, bool centre) {
    v3 p;
    float xh;
    if (LIGHT) {  // MsDir = normalize(L * off) = L
        p = mk3(fmaf(fv.L.x, s.off, px.pos.x), fmaf(fv.L.y, s.off, px.pos.y), fmaf(fv.L.z, s.off, px.pos.z));
        const float h0 = lut_fast(fv.lut, px.uc1, s.lut_v), h1 = lut_fast(fv.lut, px.uc2, s.lut_v);
        xh = __logf(fmaf(h0, h1, 1.0f));
    } else {
        p.x = fmaf(px.X.x, s.ddx, fmaf(px.Y.x, s.ddy, fmaf(px.Z.x, s.ddz, px.pos.x)));
        p.y = fmaf(px.X.y, s.ddx, fmaf(px.Y.y, s.ddy, fmaf(px.Z.y, s.ddz, px.pos.y)));
        p.z = fmaf(px.X.z, s.ddx, fmaf(px.Y.z, s.ddy, fmaf(px.Z.z, s.ddz, px.pos.z)));
        // cos between LX and the point direction: (ax*dx + ay*dy + az*dz)/|d|, ddx = dx*off
        const float c2 = fmaf(px.ax, s.ddx, fmaf(px.ay, s.ddy, px.az * s.ddz)) * s.cscale;
        const float h1 = lut_fast(fv.lut, fmaf(c2, 0.5f, 0.5f), s.lut_v);
        xh = centre ? px.xh_centre : __logf(fmaf(s.hg0, h1, 1.0f));
    }
    const float rho = tri_border(fv.pmips, s.dbase, s.dR, s.dres, p);
    const float xv = __logf(fmaf(fv.alpha64, rho, 1.0f));
    const float xs = shadow_fast(fv, s, p);
    return mk3(xv, xh, xs);
}

}  // namespace mrp
