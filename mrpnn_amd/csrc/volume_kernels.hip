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
#include "bake_order.h"
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
    float* tr_pad;       // (y, z) quads of the zero-padded, edge-replicated levels 0..6, may be null
    const float* pmips;  // zero-padded density mips 0..7, may be null (exact border sampling)
    const float4* qmips; // (y, z) quads of the padded density mips 0..7
    float alpha_lvl[8];
    long long start[9];  // flattened texel offsets of each level
    int bstart[9];       // levels 1..7 by tr_quad_block: first workgroup of each level
    v3 light;
};
// Padded TR texel value T at padded position (px, py, pz) of a level with Rp^3 padded texels, into
// the (y, z) quads that hold it (qd: the level's quads, 4 floats each): element 0 of quad
// (px, py, pz), 1 of (px, py, pz - 1), 2 of (px, py - 1, pz), 3 of (px, py - 1, pz - 1).  Every
// padded position is written, so every quad a blend reads (y, z <= Rp - 2) is complete.
__device__ inline void tr_quad_put(float* qd, int Rp, int px, int py, int pz, float T) {
    const int i = (px * Rp + py) * Rp + pz;
    qd[4 * i] = T;
    if (pz > 0) qd[4 * (i - 1) + 1] = T;
    if (py > 0) qd[4 * (i - Rp) + 2] = T;
    if (py > 0 && pz > 0) qd[4 * (i - Rp - 1) + 3] = T;
}
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
            // (y, z) quads of the padded copy with replicated edges: padded index i + 1 holds texel i, indices 0
            // and res + 1, res + 2 repeat the first and last texel of each axis, so a blend at
            // any coordinate in [0.5, res + 0.5] equals the clamp-addressed one (a blend of two
            // equal taps is exact)
            const int R = pad_dim(res);
            const int lo[3] = {x == 0 ? 0 : x + 1, y == 0 ? 0 : y + 1, z == 0 ? 0 : z + 1};
            const int hi[3] = {x == res - 1 ? res + 2 : x + 1, y == res - 1 ? res + 2 : y + 1,
                               z == res - 1 ? res + 2 : z + 1};
            float* qd = a.tr_pad + 4 * qtr_offset(m);
            for (int px = lo[0]; px <= hi[0]; ++px)
                for (int py = lo[1]; py <= hi[1]; ++py)
                    for (int pz = lo[2]; pz <= hi[2]; ++pz) tr_quad_put(qd, R, px, py, pz, T);
        }
    }
}

// The bake over the (y, z) difference quads of the zero-padded density mips, one level per
// workgroup: the level, its resolution and strides are wave-uniform (scalar registers), texel
// coordinates are shifts of the thread index, a sample's eight taps are two 16-byte loads (slices
// x and x + 1), and each thread keeps MRP_TR_BAKE_ROUND samples' loads in flight.  The wavefront's 64 lanes
// scatter over tens of cache lines per load (their rays leave the box at different distances, so
// their samples drift apart), and the texture pipeline's time follows the cache lines it
// touches: two quad loads per sample touch half the lines of four pair loads.  The arithmetic of
// every texel is k_tr_bake's (same positions, taps, blends and summation order).
#ifndef MRP_TR_BAKE_ROUND
#define MRP_TR_BAKE_ROUND 4
#endif
// TR texel i_lvl of level m (1 << 3 (7 - m) texels) over the quads.
__device__ inline void tr_quad_texel(const TrArgs& a, int m, unsigned i_lvl) {
    constexpr int K = MRP_TR_BAKE_ROUND;
    static_assert(128 % K == 0, "the round size divides the 128 steps");
    const int lres = 7 - m, res = 1 << lres;
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
            // difference quad (x, y) = (y, z), (y, z+1) - (y, z), (y+1, z), (y+1, z+1) - (y+1, z) of
            // slice x: the z lerps a + w (b - a) are one FMA each
            const float v = flerp(flerp(fmaf(wz[k], q0[k].y, q0[k].x), fmaf(wz[k], q0[k].w, q0[k].z), wy[k]),
                                  flerp(fmaf(wz[k], q1[k].y, q1[k].x), fmaf(wz[k], q1[k].w, q1[k].z), wy[k]),
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
        float* qd = a.tr_pad + 4 * qtr_offset(m);
        for (int px = lo[0]; px <= hi[0]; ++px)
            for (int py = lo[1]; py <= hi[1]; ++py)
                for (int pz = lo[2]; pz <= hi[2]; ++pz) tr_quad_put(qd, Rp, px, py, pz, T);
    }
}
// Block blk of the levels a.bstart covers (nthreads texels per workgroup), level from a.bstart.
__device__ inline void tr_quad_block(const TrArgs& a, int blk, int nthreads) {
    int m = 0;
    while (m < 7 && blk >= a.bstart[m + 1]) ++m;
    const int lres = 7 - m;
    const unsigned i_lvl = (unsigned)(blk - a.bstart[m]) * (unsigned)nthreads + threadIdx.x;
    if (i_lvl >= (1u << (3 * lres))) return;
    tr_quad_texel(a, m, i_lvl);
}

// ------------------------------------------------------------------------------------------
// TR level 0 (87 % of the bake) from LDS.  Every sample of a texel lies on the ray from the
// texel towards the light, so in sheared coordinates along the light's principal axis A
// (u' = c_U - r_U c_A, v' = c_V - r_V c_A with r = L / L_A, |r| <= 1, padded texel coordinates
// c) a texel's whole ray stays at its own (u', v').  A workgroup takes the texels of one 4 x 4
// tile of (u', v'), stages for each of the 130 padded slabs c_A = s the 10 x 10 window of density
// mip 1 that the tile's rays cross there (window origin floor(u'_lo + r_U s - |r_U| - 0.01),
// wide enough for the tile, the ray's slope over the slab pair a blend touches, the second tap
// and float slack of the running position sum), and blends every sample from LDS: the texture
// pipeline then carries no tap traffic and the concurrent march has it to itself.  Each texel's
// arithmetic is tr_quad_texel's (same positions, taps, blend and summation order), only the
// tap loads come from LDS.
// ------------------------------------------------------------------------------------------
constexpr int kL0Tile = 4;              // tile side in (u', v'), texels
constexpr int kL0Win = kL0Tile + 6;     // window side per slab
constexpr int kL0Slabs = 130;           // padded slabs that samples touch (c_A in [0.5, 128.5])
constexpr int kL0Cand = kL0Tile + 3;    // candidate texels per axis and layer
// texels per tile: 8 per axis and layer in exact arithmetic, at most 9 with rounding at the
// tile's edges
constexpr int kL0List = 128 * (kL0Tile + 1) * (kL0Tile + 1);
constexpr int kL0Threads = 512;
struct BakeL0 {
    float ru, rv;      // L_U / L_A, L_V / L_A
    float umin, vmin;  // (u', v') origin of tile (0, 0)
    int ntu, ntv;      // tiles along u', v'
    int ciu, civ;      // tile of the cube centre's (u', v'): the heaviest tiles surround it
    int nlvl;          // nlvl workgroups take levels 1..7 (a.bstart, kL0Threads texels each) ...
    int lvl_at;        // ... from rank lvl_at on, between the heavier and the lighter tiles
    int up;            // L_A > 0: rays run towards higher slabs
};
__device__ inline int l0_origin(float lo, float r, int s) { return (int)floorf(lo + r * (float)s - fabsf(r) - 0.01f); }

template <int A, int U, int V>
__global__ __launch_bounds__(kL0Threads) void k_tr_bake_l0(TrArgs a, BakeL0 b) {
    static_assert(A != U && A != V && U != V, "axis permutation");
    constexpr int WW = kL0Win * kL0Win;
    __shared__ float win[kL0Slabs * WW];
    __shared__ int base[kL0Slabs];
    __shared__ unsigned list[kL0List];
    __shared__ unsigned count;
    __shared__ int lay_lo, lay_hi;
    // levels 1..7 over the quads, in the middle of the dispatch order: at its front their
    // texture loads competed with the concurrent march's taps, at its end they were the bake's
    // last, longest workgroups
    int rank = (int)blockIdx.x;
    if (rank >= b.lvl_at && rank < b.lvl_at + b.nlvl) {
        tr_quad_block(a, rank - b.lvl_at, kL0Threads);
        return;
    }
    if (rank >= b.lvl_at + b.nlvl) rank -= b.nlvl;
    // tiles heaviest first: rings around the centre tile
    const int tile = ring_tile(rank, b.ntu, b.ntv, b.ciu, b.civ);
    if (tile < 0) return;  // not reached: ranks below ntiles map onto the grid
    const int tu = tile % b.ntu, tv = tile / b.ntu;
    const float ulo = b.umin + (float)(kL0Tile * tu), vlo = b.vmin + (float)(kL0Tile * tv);
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) {
        count = 0u;
        lay_lo = 1 << 30;
        lay_hi = -1;
    }
    __syncthreads();
    // the tile's texels: candidates around the tile per layer, kept where (u', v') falls in the
    // tile (the same float test in every workgroup, so each texel is taken exactly once)
    const int nc = 128 * kL0Cand * kL0Cand;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        const int ia = k / (kL0Cand * kL0Cand), r = k - ia * (kL0Cand * kL0Cand);
        const int du = r / kL0Cand, dv = r - du * kL0Cand;
        const float ca = (float)(ia + 1);
        const int iu = (int)floorf(fmaf(b.ru, ca, ulo)) - 2 + du;
        const int iv = (int)floorf(fmaf(b.rv, ca, vlo)) - 2 + dv;
        bool take = (unsigned)iu < 128u && (unsigned)iv < 128u;
        if (take) {
            const float up = fmaf(-b.ru, ca, (float)(iu + 1)), vp = fmaf(-b.rv, ca, (float)(iv + 1));
            take = (int)floorf((up - b.umin) * (1.0f / kL0Tile)) == tu && (int)floorf((vp - b.vmin) * (1.0f / kL0Tile)) == tv;
        }
        // active lanes are a prefix of the wavefront (k = tid + j * blockDim), so lane 0 is one
        const unsigned long long m = __ballot(take);
        unsigned at = 0u;
        if (lane == 0 && m) at = atomicAdd(&count, (unsigned)__popcll(m));
        at = __shfl(at, 0);
        const unsigned slot = at + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (take && slot < (unsigned)kL0List) list[slot] = ((unsigned)ia << 14) | ((unsigned)iu << 7) | (unsigned)iv;
        if (m) {  // the tile's layer range (a wavefront's candidates span at most two layers)
            const int l0 = __shfl(ia, __ffsll((unsigned long long)m) - 1), l1 = __shfl(ia, 63 - __clzll(m));
            if (lane == 0) {
                atomicMin(&lay_lo, l0);
                atomicMax(&lay_hi, l1);
            }
        }
    }
    __syncthreads();
    const unsigned n = min(count, (unsigned)kL0List);
    if (n == 0u) return;  // uniform: every thread read the same count after the barrier
    const int lay_lo_v = lay_lo, lay_hi_v = lay_hi;
    __syncthreads();  // every thread has read the layer range before it is reused
    // the slabs the tile's rays touch: a texel of layer l sits at c_A = l + 1 and its samples run
    // to the exit face (towards higher slabs when L_A > 0), taps at floor(c_A) and floor(c_A) + 1
    // and, of those, the slabs where the tile's window meets the padded level (elsewhere the
    // column runs outside the box, where no sample lies)
    constexpr int R = 131;
    if (threadIdx.x == 0) {
        lay_lo = 1 << 30;
        lay_hi = -1;
    }
    __syncthreads();
    {
        const int s0 = b.up ? max(lay_lo_v - 1, 0) : 0;
        const int s1 = b.up ? kL0Slabs - 1 : min(lay_hi_v + 3, kL0Slabs - 1);
        for (int s = s0 + (int)threadIdx.x; s <= s1; s += kL0Threads) {
            const int ou = l0_origin(ulo, b.ru, s), ov = l0_origin(vlo, b.rv, s);
            base[s] = s * WW - ou * kL0Win - ov;
            if (ou + kL0Win > 0 && ou < R && ov + kL0Win > 0 && ov < R) {
                atomicMin(&lay_lo, s);
                atomicMax(&lay_hi, s);
            }
        }
    }
    __syncthreads();
    const int s_lo = lay_lo, s_hi = lay_hi;  // non-empty: the tile's texels lie in the level
    // the windows (zero outside the padded level, whose own border is zero); for A = z the
    // slab index runs fastest, along the level's contiguous z rows
    const float* p1 = a.pmips + pmip_offset(1);
    const int ns = s_hi - s_lo + 1, tot = ns * WW;
    constexpr int NL = 8;  // loads in flight per thread before their LDS stores
    for (int i0 = threadIdx.x; i0 < tot; i0 += NL * kL0Threads) {
        float val[NL];
        int dst[NL];
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int i = i0 + j * kL0Threads;
            val[j] = 0.0f;
            dst[j] = -1;
            if (i < tot) {
                int s, r;
                if (A == 2) {
                    s = i % ns;
                    r = i / ns;
                } else {
                    s = i / WW;
                    r = i - s * WW;
                }
                s += s_lo;
                const int wu = r / kL0Win, wv = r - wu * kL0Win;
                int c3[3];
                c3[A] = s;
                c3[U] = l0_origin(ulo, b.ru, s) + wu;
                c3[V] = l0_origin(vlo, b.rv, s) + wv;
                const bool in = (unsigned)c3[0] < (unsigned)R && (unsigned)c3[1] < (unsigned)R && (unsigned)c3[2] < (unsigned)R;
                if (in) val[j] = p1[(c3[0] * R + c3[1]) * R + c3[2]];
                dst[j] = s * WW + r;
            }
        }
#pragma unroll
        for (int j = 0; j < NL; ++j)
            if (dst[j] >= 0) win[dst[j]] = val[j];
    }
    __syncthreads();
    constexpr int K = 4;
    for (unsigned e = threadIdx.x; e < n; e += blockDim.x) {
        const unsigned wrd = list[e];
        int c3[3];
        c3[A] = (int)(wrd >> 14);
        c3[U] = (int)((wrd >> 7) & 127u);
        c3[V] = (int)(wrd & 127u);
        const int x = c3[0], y = c3[1], z = c3[2];
        v3 uv = (mk3((float)x, (float)y, (float)z) + 0.5f) / 128.0f;
        v3 ori = uv - 0.5f;
        float dis = ray_box_distance(ori, a.light);
        float st = dis / 128.0f;
        v3 lp = ori;
        float sum = 0.0f;
        for (int i = 0; i < 128; i += K) {
            float wx[K], wy[K], wz[K];
            int a0[K], a1[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                lp = lp + a.light * st;
                const v3 c = padded_coord(lp, 128.0f);
                wx[k] = __builtin_amdgcn_fractf(c.x);
                wy[k] = __builtin_amdgcn_fractf(c.y);
                wz[k] = __builtin_amdgcn_fractf(c.z);
                // non-negative coordinates: truncation is the floor
                int f3[3] = {(int)c.x, (int)c.y, (int)c.z};
                const int off = (int)mad_u24_v((unsigned)f3[U], (unsigned)kL0Win, (unsigned)f3[V]);
                a0[k] = base[f3[A]] + off;
                a1[k] = base[f3[A] + 1] + off;
            }
            float t[K][2][2][2];  // [sample][slab +0/+1][u +0/+1][v +0/+1]
#pragma unroll
            for (int k = 0; k < K; ++k) {
                t[k][0][0][0] = win[a0[k]];
                t[k][0][0][1] = win[a0[k] + 1];
                t[k][0][1][0] = win[a0[k] + kL0Win];
                t[k][0][1][1] = win[a0[k] + kL0Win + 1];
                t[k][1][0][0] = win[a1[k]];
                t[k][1][0][1] = win[a1[k] + 1];
                t[k][1][1][0] = win[a1[k] + kL0Win];
                t[k][1][1][1] = win[a1[k] + kL0Win + 1];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                // corner (dx, dy, dz) of the blend is window tap [d_A][d_U][d_V]
                float cc[2][2][2];
#pragma unroll
                for (int dx = 0; dx < 2; ++dx)
#pragma unroll
                    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                        for (int dz = 0; dz < 2; ++dz) {
                            const int d[3] = {dx, dy, dz};
                            cc[dx][dy][dz] = t[k][d[A]][d[U]][d[V]];
                        }
                const float v = flerp(flerp(flerp(cc[0][0][0], cc[0][0][1], wz[k]), flerp(cc[0][1][0], cc[0][1][1], wz[k]), wy[k]),
                                      flerp(flerp(cc[1][0][0], cc[1][0][1], wz[k]), flerp(cc[1][1][0], cc[1][1][1], wz[k]), wy[k]),
                                      wx[k]);
                sum = sum + v;
            }
        }
        const float T = expf(-sum * a.alpha_lvl[0] * st) * 1.0f;
        a.tr[0][(x * 128 + y) * 128 + z] = T;
        if (a.tr_pad) {  // quads of the padded level with replicated edges (as k_tr_bake)
            const int Rp = pad_dim(128);
            const int lo[3] = {x == 0 ? 0 : x + 1, y == 0 ? 0 : y + 1, z == 0 ? 0 : z + 1};
            const int hi[3] = {x == 127 ? 130 : x + 1, y == 127 ? 130 : y + 1, z == 127 ? 130 : z + 1};
            for (int px = lo[0]; px <= hi[0]; ++px)
                for (int py = lo[1]; py <= hi[1]; ++py)
                    for (int pz = lo[2]; pz <= hi[2]; ++pz) tr_quad_put(a.tr_pad, Rp, px, py, pz, T);
        }
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

// (y, z) difference quads of the padded density mips 0..7 (every padded texel, the border
// included; taps beyond the padded level read 0).
__global__ __launch_bounds__(256) void k_quad_mips(const float* __restrict__ pm, float4* __restrict__ q) {
    const long long total = kQuadTexels;
    for (long long g = (long long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long long)gridDim.x * 256) {
        int m = 0;
        while (m < 7 && g >= qmip_offset(m + 1)) ++m;
        const int R = pad_dim(256 >> m);
        const int l = (int)(g - qmip_offset(m));
        const int z = l % R, y = (l / R) % R;
        const float* p = pm + pmip_offset(m);
        const bool zn = z + 1 < R, yn = y + 1 < R;
        // difference quads: the z neighbours as differences, rounded as a lerp's b - a is
        float4 o;
        o.x = p[l];
        o.y = (zn ? p[l + 1] : 0.0f) - o.x;
        o.z = yn ? p[l + R] : 0.0f;
        o.w = ((yn && zn) ? p[l + R + 1] : 0.0f) - o.z;
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
        // level 0 from LDS along the light's principal axis
        const float Lc[3] = {light.x, light.y, light.z};
        int A = 0;
        for (int i = 1; i < 3; ++i)
            if (fabsf(Lc[i]) > fabsf(Lc[A])) A = i;
        const int U = A == 0 ? 1 : 0, Vx = A == 2 ? 1 : 2;
        BakeL0 b;
        b.ru = Lc[U] / Lc[A];
        b.rv = Lc[Vx] / Lc[A];
        float lo[2] = {1e30f, 1e30f}, hi[2] = {-1e30f, -1e30f};
        for (float ca : {1.0f, 128.0f})
            for (float cw : {1.0f, 128.0f}) {
                const float r2[2] = {b.ru, b.rv};
                for (int j = 0; j < 2; ++j) {
                    const float w = cw - r2[j] * ca;
                    lo[j] = fminf(lo[j], w);
                    hi[j] = fmaxf(hi[j], w);
                }
            }
        b.umin = floorf(lo[0]) - 1.0f;
        b.vmin = floorf(lo[1]) - 1.0f;
        b.ntu = (int)((hi[0] + 1.0f - b.umin) / kL0Tile) + 1;
        b.ntv = (int)((hi[1] + 1.0f - b.vmin) / kL0Tile) + 1;
        // the cube centre (padded coordinate 64.5 on every axis) in (u', v') tiles
        const float cu = (64.5f - b.ru * 64.5f - b.umin) / kL0Tile, cv = (64.5f - b.rv * 64.5f - b.vmin) / kL0Tile;
        b.ciu = std::min(std::max((int)floorf(cu), 0), b.ntu - 1);
        b.civ = std::min(std::max((int)floorf(cv), 0), b.ntv - 1);
        b.up = Lc[A] > 0.0f ? 1 : 0;
        // levels 1..7 by the quad path in workgroups of their own (level 0 gets none)
        int nb = 0;
        a.bstart[0] = 0;
        for (int m = 1; m < 8; ++m) {
            a.bstart[m] = nb;
            nb += (int)(((128 >> m) * (128 >> m) * (128 >> m) + kL0Threads - 1) / kL0Threads);
        }
        a.bstart[8] = nb;
        b.nlvl = nb;
        b.lvl_at = b.ntu * b.ntv / 2;
        const dim3 g0(b.ntu * b.ntv + nb);
        if (A == 0) hipLaunchKernelGGL((k_tr_bake_l0<0, 1, 2>), g0, dim3(kL0Threads), 0, s, a, b);
        else if (A == 1) hipLaunchKernelGGL((k_tr_bake_l0<1, 0, 2>), g0, dim3(kL0Threads), 0, s, a, b);
        else hipLaunchKernelGGL((k_tr_bake_l0<2, 0, 1>), g0, dim3(kL0Threads), 0, s, a, b);
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
