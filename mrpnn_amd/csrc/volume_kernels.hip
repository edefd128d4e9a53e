// Per-volume and per-light preparation kernels:
//   * seeded fBm cloud generation (benchmark inputs);
//   * density mip pyramid, VolumeRender::Update (core/volume.cu:1443-1531), moved from the
//     reference's OpenMP loop to the GPU: exact 2x levels are 2x2x2 box means with an FP64
//     sum, other ratios are the zero-border trilinear resample of the host Sample()
//     (core/volume.cu:845-863) at cell centres;
//   * transmittance mips, Fill_TR / Update_TR (core/volume.cu:1535-1560, 1589-1667), all
//     eight levels in one launch;
//   * HG LUT, Fill_Hg (core/volume.cu:1322-1400).
// All are HBM-streaming or ALU kernels with coalesced z-fastest access.
#include "dev_common.hpp"
#include "fbm.h"
#include "kernels.h"
#include "ref_tables.h"

namespace mrp {

// ------------------------------------------------------------------------------------------
// fBm generator: raw values, grid min/max (order-preserving uint atomics), normalise.
// ------------------------------------------------------------------------------------------
__device__ inline uint32_t f2ord(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

__global__ __launch_bounds__(256) void k_fbm_raw(float* __restrict__ out, int N, int octaves, float base_freq,
                                                 const int* __restrict__ perm_g, uint32_t* __restrict__ minmax) {
    __shared__ int perm[512];
    for (int i = threadIdx.x; i < 512; i += 256) perm[i] = perm_g[i];
    __syncthreads();
    const size_t total = (size_t)N * N * N;
    float lo = 3.0e38f, hi = -3.0e38f;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        int z = (int)(idx % N);
        int y = (int)((idx / N) % N);
        int x = (int)(idx / ((size_t)N * N));
        float v = fbm_voxel(perm, N, octaves, base_freq, x, y, z);
        out[idx] = v;
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    // wavefront reduction then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&minmax[0], f2ord(lo));
        atomicMax(&minmax[1], f2ord(hi));
    }
}

__global__ __launch_bounds__(256) void k_fbm_normalize(float* __restrict__ v, size_t total,
                                                       const uint32_t* __restrict__ minmax) {
    const float lo = ord2f(minmax[0]), hi = ord2f(minmax[1]);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256)
        v[i] = fbm_density(v[i], lo, hi);
}

// ------------------------------------------------------------------------------------------
// Pyramid
// ------------------------------------------------------------------------------------------
// Exact 2x: dst[x][y][z] = mean of the 2x2x2 children (each clamped >= 0), summed in FP64.
// The reference's trilinear at (2i+0.5) reduces to this mean with seven fp32 roundings; the
// FP64 sum rounds once.
__global__ __launch_bounds__(256) void k_box2(const float* __restrict__ src, float* __restrict__ dst, int res) {
    const size_t total = (size_t)res * res * res;
    const int sres = res * 2;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        int z = (int)(idx % res);
        int y = (int)((idx / res) % res);
        int x = (int)(idx / ((size_t)res * res));
        double s = 0.0;
#pragma unroll
        for (int dx = 0; dx < 2; ++dx)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                const float2 p =
                    *(const float2*)(src + (((size_t)(2 * x + dx) * sres + (2 * y + dy)) * sres + 2 * z));
                s += (double)fmaxf(0.0f, p.x);
                s += (double)fmaxf(0.0f, p.y);
            }
        dst[idx] = (float)(s * 0.125);
    }
}

// General ratio: Sample(source, sres, uvw) at cell centres, zero border, max(0,.) taps
// (ParallelFill omp.cpp:4-18 computes u = (i+0.5)/res in double).
__device__ inline float fetch_b0(const float* __restrict__ d, int res, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return 0.0f;
    return fmaxf(0.0f, d[((size_t)x * res + y) * res + z]);
}
__global__ __launch_bounds__(256) void k_resample(const float* __restrict__ src, int sres, float* __restrict__ dst,
                                                  int res) {
    const size_t total = (size_t)res * res * res;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
        int k = (int)(idx % res);
        int j = (int)((idx / res) % res);
        int i = (int)(idx / ((size_t)res * res));
        v3 uv = mk3((float)((i + 0.5) / res), (float)((j + 0.5) / res), (float)((k + 0.5) / res));
        v3 pos = uv * (float)sres - 0.5f;
        int x = (int)floorf(pos.x), y = (int)floorf(pos.y), z = (int)floorf(pos.z);
        v3 w = pos - mk3((float)x, (float)y, (float)z);
        const float* d = src;
        float r = lerp_ref(lerp_ref(lerp_ref(fetch_b0(d, sres, x, y, z), fetch_b0(d, sres, x, y, z + 1), w.z),
                                    lerp_ref(fetch_b0(d, sres, x, y + 1, z), fetch_b0(d, sres, x, y + 1, z + 1), w.z),
                                    w.y),
                           lerp_ref(lerp_ref(fetch_b0(d, sres, x + 1, y, z), fetch_b0(d, sres, x + 1, y, z + 1), w.z),
                                    lerp_ref(fetch_b0(d, sres, x + 1, y + 1, z),
                                             fetch_b0(d, sres, x + 1, y + 1, z + 1), w.z),
                                    w.y),
                           w.x);
        dst[idx] = r;
    }
}

// ------------------------------------------------------------------------------------------
// Transmittance mips: level m (res 128>>m) marches 128 steps towards the light through
// density mip m+1 with alpha / 1.73^(m+1) (Fill_TR, volume.cu:1535-1560).
// ------------------------------------------------------------------------------------------
struct TrArgs {
    const float* mips[9];
    float* tr[8];
    float* tr_pad;       // zero-padded copy of levels 0..6 (fast gather), may be null
    const float* pmips;  // zero-padded density mips 0..7, may be null (exact border sampling)
    const float4* qmips; // (y, z) quads of the padded mips 1..7 (k_tr_bake_lvl)
    float alpha_lvl[8];
    long long start[9];  // flattened texel offsets of each level
    int bstart[9];       // k_tr_bake_lvl: first workgroup of each level
    v3 light;
};
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_tr_bake(TrArgs a) {
    const long long total = a.start[8];
    for (long long g = (long long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long long)gridDim.x * 256) {
        int m = 0;
        while (m < 7 && g >= a.start[m + 1]) ++m;
        const int res = 128 >> m;
        const long long i_lvl = g - a.start[m];
        const int z = (int)(i_lvl % res);
        const int y = (int)((i_lvl / res) % res);
        const int x = (int)(i_lvl / ((long long)res * res));
        const long long l = i_lvl;
        v3 uv = (mk3((float)x, (float)y, (float)z) + 0.5f) / (float)res;
        v3 ori = uv - 0.5f;
        float dis = ray_box_distance(ori, a.light);
        float st = dis / 128.0f;
        v3 lp = ori;
        float sum = 0.0f;
        const int sres = 256 >> (m + 1);
        if (a.pmips && m + 1 < 8) {
            // zero-padded copy of the density mip: no per-tap border tests.  Every sample lies
            // on the segment from the texel centre to the box exit, so its padded coordinate
            // is in [0.5, sres + 0.5] and both taps per axis are inside the padded level:
            // the zero-border clamp of tri_border_lvl never changes a coordinate and is left out
            const unsigned base = (unsigned)pmip_offset(m + 1) * 4u;
            const float fres = (float)sres;
            const float R = fres + 3.0f;
            const unsigned Ru = (unsigned)R;
            const unsigned R4 = Ru * 4u, RR4 = Ru * Ru * 4u;
            const BufRsrc rs = make_rsrc(a.pmips, kPadMipFloats);
            // four samples per round: positions (the reference's running sum), then all 16
            // tap-pair loads, then the blends, added to the sum in step order
            for (int i = 0; i < 128; i += 4) {
                float wx[4], wy[4], wz[4];
                int o[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    lp = lp + a.light * st;
                    const v3 c = padded_coord(lp, fres);
                    wx[k] = __builtin_amdgcn_fractf(c.x);
                    wy[k] = __builtin_amdgcn_fractf(c.y);
                    wz[k] = __builtin_amdgcn_fractf(c.z);
                    // source levels 1..7 (at most 131^3 padded texels < 2^24): exact in float
                    const float lin = fmaf(fmaf(floorf(c.x), R, floorf(c.y)), R, floorf(c.z));
                    o[k] = (int)((unsigned)lin * 4u + base);
                }
                f32x2_t t[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    t[k][0] = ld_pair(rs, o[k]);
                    t[k][1] = ld_pair(rs, o[k] + (int)R4);
                    t[k][2] = ld_pair(rs, o[k] + (int)RR4);
                    t[k][3] = ld_pair(rs, o[k] + (int)(RR4 + R4));
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float v = flerp(flerp(flerp(t[k][0].x, t[k][0].y, wz[k]), flerp(t[k][1].x, t[k][1].y, wz[k]), wy[k]),
                                          flerp(flerp(t[k][2].x, t[k][2].y, wz[k]), flerp(t[k][3].x, t[k][3].y, wz[k]), wy[k]),
                                          wx[k]);
                    sum = sum + v;
                }
            }
        } else {
            const float* src = a.mips[m + 1];
            for (int i = 0; i < 128; ++i) {
                lp = lp + a.light * st;
                sum = sum + sample_border_exact(src, sres, lp + 0.5f);
            }
        }
        const float T = expf(-sum * a.alpha_lvl[m] * st) * 1.0f;
        a.tr[m][l] = T;
        if (a.tr_pad && m < 7) {
            // padded copy with replicated edges: padded index i + 1 holds texel i, indices 0
            // and res + 1, res + 2 repeat the first and last texel of each axis, so a blend at
            // any coordinate in [0.5, res + 0.5] equals the clamp-addressed one (a blend of two
            // equal taps is exact)
            const int R = pad_dim(res);
            const int lo[3] = {x == 0 ? 0 : x + 1, y == 0 ? 0 : y + 1, z == 0 ? 0 : z + 1};
            const int hi[3] = {x == res - 1 ? res + 2 : x + 1, y == res - 1 ? res + 2 : y + 1,
                               z == res - 1 ? res + 2 : z + 1};
            float* dst = a.tr_pad + ptr_offset(m);
            for (int px = lo[0]; px <= hi[0]; ++px)
                for (int py = lo[1]; py <= hi[1]; ++py)
                    for (int pz = lo[2]; pz <= hi[2]; ++pz) dst[((long long)px * R + py) * R + pz] = T;
        }
    }
}

// The bake over the (y, z) quads of the zero-padded density mips, one level per workgroup: the
// level, its resolution and strides are wave-uniform (scalar registers), texel coordinates are
// shifts of the thread index, a sample's eight taps are two 16-byte loads (slices x and x + 1),
// and each thread keeps MRP_TR_BAKE_ROUND samples' loads in flight.  The wavefront's 64 lanes
// scatter over tens of cache lines per load (their rays leave the box at different distances, so
// their samples drift apart), and the texture pipeline's time follows the cache lines it
// touches: two quad loads per sample touch half the lines of four pair loads.  The arithmetic of
// every texel is k_tr_bake's (same positions, taps, blends and summation order).
#ifndef MRP_TR_BAKE_ROUND
#define MRP_TR_BAKE_ROUND 4
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_tr_bake_lvl(TrArgs a) {
    constexpr int K = MRP_TR_BAKE_ROUND;
    static_assert(128 % K == 0, "the round size divides the 128 steps");
    int m = 0;
    while (m < 7 && (int)blockIdx.x >= a.bstart[m + 1]) ++m;
    const int lres = 7 - m, res = 1 << lres;
    const unsigned i_lvl = (unsigned)((int)blockIdx.x - a.bstart[m]) * 256u + threadIdx.x;
    if (i_lvl >= (1u << (3 * lres))) return;
    const int z = (int)(i_lvl & (unsigned)(res - 1));
    const int y = (int)((i_lvl >> lres) & (unsigned)(res - 1));
    const int x = (int)(i_lvl >> (2 * lres));
    v3 uv = (mk3((float)x, (float)y, (float)z) + 0.5f) / (float)res;
    v3 ori = uv - 0.5f;
    float dis = ray_box_distance(ori, a.light);
    float st = dis / 128.0f;
    v3 lp = ori;
    float sum = 0.0f;
    const int sres = 256 >> (m + 1);
    if (m + 1 >= 8) {  // level 7 samples the 1^3 mip 8, which has no padded copy
        const float* src = a.mips[m + 1];
        for (int i = 0; i < 128; ++i) {
            lp = lp + a.light * st;
            sum = sum + sample_border_exact(src, sres, lp + 0.5f);
        }
        a.tr[m][i_lvl] = expf(-sum * a.alpha_lvl[m] * st) * 1.0f;
        return;
    }
    const unsigned base = (unsigned)qmip_offset(m + 1) * 16u;
    const float fres = (float)sres;
    const float R = fres + 3.0f;
    const unsigned Ru = (unsigned)R;
    const unsigned RR16 = Ru * Ru * 16u;
    const BufRsrc rs = __builtin_amdgcn_make_buffer_rsrc((void*)a.qmips, (short)0, (int)(kQuadTexels * 16), 0x00020000);
    for (int i = 0; i < 128; i += K) {
        float wx[K], wy[K], wz[K];
        int o[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            lp = lp + a.light * st;
            const v3 c = padded_coord(lp, fres);
            wx[k] = __builtin_amdgcn_fractf(c.x);
            wy[k] = __builtin_amdgcn_fractf(c.y);
            wz[k] = __builtin_amdgcn_fractf(c.z);
            // source levels 1..7 (at most 131^3 padded texels < 2^24): exact in float
            const float lin = fmaf(fmaf(floorf(c.x), R, floorf(c.y)), R, floorf(c.z));
            o[k] = (int)((unsigned)lin * 16u + base);
        }
        f32x4_t q0[K], q1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            q0[k] = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(rs, o[k], 0, 0));
            q1[k] = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(rs, o[k] + (int)RR16, 0, 0));
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // quad (x, y) = (y, z), (y, z+1), (y+1, z), (y+1, z+1) of slice x
            const float v = flerp(flerp(flerp(q0[k].x, q0[k].y, wz[k]), flerp(q0[k].z, q0[k].w, wz[k]), wy[k]),
                                  flerp(flerp(q1[k].x, q1[k].y, wz[k]), flerp(q1[k].z, q1[k].w, wz[k]), wy[k]),
                                  wx[k]);
            sum = sum + v;
        }
    }
    const float T = expf(-sum * a.alpha_lvl[m] * st) * 1.0f;
    a.tr[m][i_lvl] = T;
    if (a.tr_pad && m < 7) {
        // padded copy with replicated edges (as k_tr_bake)
        const int Rp = pad_dim(res);
        const int lo[3] = {x == 0 ? 0 : x + 1, y == 0 ? 0 : y + 1, z == 0 ? 0 : z + 1};
        const int hi[3] = {x == res - 1 ? res + 2 : x + 1, y == res - 1 ? res + 2 : y + 1,
                           z == res - 1 ? res + 2 : z + 1};
        float* dst = a.tr_pad + ptr_offset(m);
        for (int px = lo[0]; px <= hi[0]; ++px)
            for (int py = lo[1]; py <= hi[1]; ++py)
                for (int pz = lo[2]; pz <= hi[2]; ++pz) dst[(px * Rp + py) * Rp + pz] = T;
    }
}

// ------------------------------------------------------------------------------------------
// HG LUT (Fill_Hg, volume.cu:1322-1400): texel (i = cos index, j = cone-angle index).
// ------------------------------------------------------------------------------------------
__constant__ unsigned c_halton_primes[5] = MRP_HALTON_PRIMES_INIT;
__constant__ unsigned c_halton_sums[5] = MRP_HALTON_PRIME_SUMS_INIT;
__constant__ unsigned c_halton_perm[28] = MRP_HALTON_PERM_INIT;

__device__ inline float halton_scrambled(unsigned dim, unsigned a) {  // volume.cu:1300-1316
    const unsigned base = c_halton_primes[dim], offset = c_halton_sums[dim];
    const float inv_base = 1.0f / (float)base;
    unsigned rev = 0;
    float inv_base_n = 1.0f;
    while (a) {
        unsigned next = a / base;
        unsigned digit = a - next * base;
        rev = rev * base + c_halton_perm[offset + digit];
        inv_base_n *= inv_base;
        a = next;
    }
    float r = inv_base_n * ((float)rev + inv_base * (float)c_halton_perm[offset] / (1.0f - inv_base));
    return fminf(r, 0.99999994f);
}
__device__ inline float frac_ref(float f) { return f - (float)(long)f; }  // vector.cu:20

__global__ __launch_bounds__(64) void k_hg_lut(float* __restrict__ lut, float g, uint32_t seed, float mis_a) {
    const int i = blockIdx.x * 8 + (threadIdx.x & 7);
    const int j = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (i >= 512 || j >= 512) return;
    // texel centres (i + 0.5) / 512 and their cos are exact in float; the cone angle is
    // aa * pi/3 (volume.cu:1330-1332 in double; one float rounding here)
    const float costheta = ((float)i + 0.5f) / 512.0f * 2.0f - 1.0f;
    const float aa = ((float)j + 0.5f) / 512.0f;
    const float angle = aa * 1.04719755f;
    PhiloxStream rs(seed, 0u, (uint32_t)(i + j * 512), RNG_LUT, 0u);
    float bn = rs.next();
    float hg = 0.0f;
    float R = tanf(angle);
    const v3 view = mk3(1, 0, 0);
    v3 D = mk3(costheta, sqrtf(1.0f - costheta * costheta), 0.0f);
    float mis = 0.0f;
    {
        float p = R * R - 1.0f + D.x * D.x;
        if (p > 0.0f) {
            float sdotl = g > 0.0f ? D.x : -D.x;
            float w = sqrtf(p);
            float d0 = fmaxf(0.0f, sdotl - w) / R;
            float d1 = (sdotl + w) / R;
            float d = d1 * d1 * d1 - d0 * d0 * d0;
            mis = (float)(d1 > d0) * d;
        }
        mis = fminf(fmaxf(mis, 1e-8f), 1.0f);
        mis = 1.0f - 1.0f / (1.0f + mis_a * mis / (1.0f - mis_a));
    }
    for (int k = 0; k < 1024; ++k) {
        if (frac_ref(halton_scrambled(4, k) + bn) < mis) {
            float rx = frac_ref(halton_scrambled(0, k) + bn), ry = frac_ref(halton_scrambled(1, k) + bn);
            v3 l = sample_hg(rx, ry, view, g);
            float sdotl = dot3(D, l);
            float p = R * R - 1.0f + sdotl * sdotl;
            if (p <= 0.0f) continue;
            float w = sqrtf(p);
            float d0 = sdotl - w;
            float d1 = (sdotl + w) / R;
            d0 = fmaxf(0.0f, d0) / R;
            float d = d1 * d1 * d1 - d0 * d0 * d0;
            hg += fmaxf(0.0f, d / 4.0f);
        } else {
            float rx = frac_ref(halton_scrambled(0, k) + bn), ry = frac_ref(halton_scrambled(1, k) + bn),
                  rz = frac_ref(halton_scrambled(2, k) + bn);
            const float theta = rx * 6.28318531f;
            const float phi = acosf(2.0f * ry - 1.0f);
            float r = powf(rz, 1.0f / 3.0f);
            float st = sinf(theta), ct = cosf(theta), sp = sinf(phi), cp = cosf(phi);
            float x = R * r * sp * ct, y = R * r * sp * st, z = R * r * cp;
            v3 rp = mk3(x, y, z) + D;
            float c = dot3(normalize3(rp), view);
            hg += henyey_greenstein(c, g);
        }
    }
    hg /= 1024.0f;
    lut[j * 512 + i] = hg;
}

// ------------------------------------------------------------------------------------------
// Zero-padded copies of density mips 0..7 for the fast gather (borders stay zero from the
// allocation-time clear).
struct PadArgs {
    const float* mips[8];
    long long start[9];
};
__global__ __launch_bounds__(256) void k_pad_mips(PadArgs a, float* __restrict__ pm) {
    const long long total = a.start[8];
    for (long long g = (long long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long long)gridDim.x * 256) {
        int m = 0;
        while (m < 7 && g >= a.start[m + 1]) ++m;
        const int res = 256 >> m, R = pad_dim(res);
        const long long l = g - a.start[m];
        const int z = (int)(l % res), y = (int)((l / res) % res), x = (int)(l / ((long long)res * res));
        pm[pmip_offset(m) + ((long long)(x + 1) * R + (y + 1)) * R + (z + 1)] = a.mips[m][l];
    }
}

// (y, z) quads of the padded density mips 1..7 (every padded texel, the border included; taps
// beyond the padded level read 0).
__global__ __launch_bounds__(256) void k_quad_mips(const float* __restrict__ pm, float4* __restrict__ q) {
    const long long total = kQuadTexels;
    for (long long g = (long long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long long)gridDim.x * 256) {
        int m = 1;
        while (m < 7 && g >= qmip_offset(m + 1)) ++m;
        const int R = pad_dim(256 >> m);
        const int l = (int)(g - qmip_offset(m));
        const int z = l % R, y = (l / R) % R;
        const float* p = pm + pmip_offset(m);
        const bool zn = z + 1 < R, yn = y + 1 < R;
        float4 o;
        o.x = p[l];
        o.y = zn ? p[l + 1] : 0.0f;
        o.z = yn ? p[l + R] : 0.0f;
        o.w = (yn && zn) ? p[l + R + 1] : 0.0f;
        q[g] = o;
    }
}

// Brick emptiness of the full-resolution grid for the first-hit march: brick b covers
// texels [8b-1, 8b+7] per axis (the corners a sample in that brick can read); 1 = all zero.
__global__ __launch_bounds__(256) void k_occupancy(const float* __restrict__ grid, int N, uint8_t* __restrict__ occ,
                                                   int n) {
    const long long total = (long long)n * n * n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        const int bz = (int)(i % n), by = (int)((i / n) % n), bx = (int)(i / ((long long)n * n));
        const int x0 = max(8 * bx - 1, 0), x1 = min(8 * bx + 7, N - 1);
        const int y0 = max(8 * by - 1, 0), y1 = min(8 * by + 7, N - 1);
        const int z0 = max(8 * bz - 1, 0), z1 = min(8 * bz + 7, N - 1);
        bool empty = true;
        for (int x = x0; x <= x1 && empty; ++x)
            for (int y = y0; y <= y1 && empty; ++y)
                for (int z = z0; z <= z1; ++z)
                    if (grid[((size_t)x * N + y) * N + z] != 0.0f) {
                        empty = false;
                        break;
                    }
        occ[i] = empty ? 1 : 0;
    }
}

// Fast-gather stencil table: direction * offset, LUT v, the per-point LUT lookup of sphere
// points (dot(XMain, direction) = dx/|d| for an orthonormal basis), padded level geometry.
__global__ void k_fast_stencil(const StencilEntry* __restrict__ st, const float* __restrict__ lut,
                               FastStencil* __restrict__ out) {
    const int i = threadIdx.x;
    if (i >= kPC) return;
    const StencilEntry e = st[i];
    FastStencil f;
    f.kind = i == 0 ? 0 : (e.type >= 4 ? 1 : 2);
    f.ddx = e.dx * e.offset;
    f.ddy = e.dy * e.offset;
    f.ddz = e.dz * e.offset;
    const float len = sqrtf(e.dx * e.dx + e.dy * e.dy + e.dz * e.dz);
    f.cscale = f.kind == 1 ? 1.0f / (len * e.offset) : 0.0f;
    f.off = e.offset;
    f.lut_v = e.lut_v;
    f.hg0 = f.kind == 1 ? sample_lut_exact(lut, (e.dx / len) * 0.5f + 0.5f, e.lut_v) : 0.0f;
    const int tm = e.mip - 1 > 0 ? e.mip - 1 : 0;
    const unsigned dR = (unsigned)pad_dim(256 >> e.mip), tR = (unsigned)pad_dim(128 >> tm);
    f.dres = (float)(256 >> e.mip);
    f.tres = (float)(128 >> tm);
    f.dRu = dR;
    f.tRu = tR;
    f.dbase = (unsigned)pmip_offset(e.mip) * 4u;
    f.tbase = (unsigned)ptr_offset(tm) * 4u;
    f.dR4 = dR * 4u;
    f.dRR4 = dR * dR * 4u;
    f.tR4 = tR * 4u;
    f.tRR4 = tR * tR * 4u;
    f.lbase = (unsigned)(i * kLutRow * 4);
    f.pad0 = 0;
    out[i] = f;
}

// LUT rows of the fast gather: row i = the LUT interpolated along v at stencil point i's v.
__global__ __launch_bounds__(512) void k_lut_rows(const StencilEntry* __restrict__ st, const float* __restrict__ lut,
                                                  float* __restrict__ rows) {
    const int i = blockIdx.x, x = threadIdx.x;
    const float v = st[i].lut_v;
    const float py = fminf(fmaxf(fmaf(v, 512.0f, -0.5f), 0.0f), 511.0f);
    const unsigned y0 = (unsigned)py, y1 = min(y0 + 1u, 511u);
    const float wy = __builtin_amdgcn_fractf(py);
    const float r = flerp(lut[y0 * 512u + x], lut[y1 * 512u + x], wy);
    rows[i * kLutRow + x] = r;
    if (x == 511)
        for (int k = 512; k < kLutRow; ++k) rows[i * kLutRow + k] = r;
}

static inline int grid_for(size_t total, int block) {
    size_t b = (total + block - 1) / block;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

hipError_t launch_fbm(float* out, int N, int octaves, float base_freq, const int* perm, uint32_t* minmax,
                      hipStream_t s) {
    const size_t total = (size_t)N * N * N;
    hipLaunchKernelGGL(k_fbm_raw, dim3(grid_for(total, 256)), dim3(256), 0, s, out, N, octaves, base_freq, perm,
                       minmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fbm_normalize, dim3(grid_for(total, 256)), dim3(256), 0, s, out, total,
                       (const uint32_t*)minmax);
    return hipGetLastError();
}
hipError_t launch_box2(const float* src, float* dst, int res, hipStream_t s) {
    hipLaunchKernelGGL(k_box2, dim3(grid_for((size_t)res * res * res, 256)), dim3(256), 0, s, src, dst, res);
    return hipGetLastError();
}
hipError_t launch_resample(const float* src, int sres, float* dst, int res, hipStream_t s) {
    hipLaunchKernelGGL(k_resample, dim3(grid_for((size_t)res * res * res, 256)), dim3(256), 0, s, src, sres, dst,
                       res);
    return hipGetLastError();
}
hipError_t launch_tr_bake(const float* const mips[9], const float* pmips, const float4* qmips, float* const tr[8],
                          float* tr_pad,
                          const float alpha_lvl[8], v3 light, hipStream_t s) {
    TrArgs a;
    a.tr_pad = tr_pad;
    a.pmips = pmips;
    long long acc = 0;
    for (int m = 0; m < 9; ++m) a.mips[m] = mips[m];
    for (int m = 0; m < 8; ++m) {
        a.tr[m] = tr[m];
        a.alpha_lvl[m] = alpha_lvl[m];
        a.start[m] = acc;
        acc += (long long)(128 >> m) * (128 >> m) * (128 >> m);
    }
    a.start[8] = acc;
    a.light = light;
    a.qmips = qmips;
    if (pmips && qmips) {
        int nb = 0;
        for (int m = 0; m < 8; ++m) {
            a.bstart[m] = nb;
            nb += (int)(((128 >> m) * (128 >> m) * (128 >> m) + 255) / 256);
        }
        a.bstart[8] = nb;
        hipLaunchKernelGGL(k_tr_bake_lvl, dim3(nb), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_tr_bake, dim3(grid_for((size_t)acc, 256)), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}
hipError_t launch_pad_mips(const float* const mips[9], float* pmips, hipStream_t s) {
    PadArgs a;
    long long acc = 0;
    for (int m = 0; m < 8; ++m) {
        a.mips[m] = mips[m];
        a.start[m] = acc;
        acc += (long long)(256 >> m) * (256 >> m) * (256 >> m);
    }
    a.start[8] = acc;
    hipLaunchKernelGGL(k_pad_mips, dim3(grid_for((size_t)acc, 256)), dim3(256), 0, s, a, pmips);
    return hipGetLastError();
}
hipError_t launch_quad_mips(const float* pmips, float4* qmips, hipStream_t s) {
    hipLaunchKernelGGL(k_quad_mips, dim3(grid_for((size_t)kQuadTexels, 256)), dim3(256), 0, s, pmips, qmips);
    return hipGetLastError();
}
hipError_t launch_occupancy(const float* grid, int N, uint8_t* occ, int occ_n, hipStream_t s) {
    hipLaunchKernelGGL(k_occupancy, dim3(grid_for((size_t)occ_n * occ_n * occ_n, 256)), dim3(256), 0, s, grid, N, occ,
                       occ_n);
    return hipGetLastError();
}
hipError_t launch_fast_stencil(const StencilEntry* st, const float* lut, void* fast, float* lut_rows, hipStream_t s) {
    hipLaunchKernelGGL(k_fast_stencil, dim3(1), dim3(kPC), 0, s, st, lut, (FastStencil*)fast);
    hipLaunchKernelGGL(k_lut_rows, dim3(kPC), dim3(512), 0, s, st, lut, lut_rows);
    return hipGetLastError();
}

hipError_t launch_hg_lut(float* lut, float g, uint32_t seed, hipStream_t s) {
    // anisotropy weight of the mixture (volume.cu:1357-1358): |g| in float, the rest in double
    const double a = (double)fabsf(g);
    const float mis_a = a < 0.6 ? 0.0f : (a < 0.7 ? (float)((a - 0.6) * 9) : (float)(pow((a - 0.7) / 0.3, 0.05) * 0.1 + 0.9));
    hipLaunchKernelGGL(k_hg_lut, dim3(64, 64), dim3(64), 0, s, lut, g, seed, mis_a);
    return hipGetLastError();
}

}  // namespace mrp
