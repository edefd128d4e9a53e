// Seeded fBm cloud generator: the synthetic density volumes of the benchmark workloads.
//
// Improved gradient noise (K. Perlin, "Improving Noise", SIGGRAPH 2002: quintic fade and
// 12 edge gradients) summed over octaves at lattice-cell centres (i+1/2)/N, base frequency
// f0, gain 0.5, lacunarity 2.  The permutation is a Fisher-Yates shuffle driven by Philox
// keyed on the seed.  The raw sum is later normalised by its grid min/max, values below
// 0.4 become 0 and the rest map to (u-0.4)/0.6, stored in the reference's (x*N+y)*N+z
// layout (core/volume.cu:1179).  The reference's own noise (core/vector.cu:408-495) is
// never called on its render path and is not followed.
//
// Only +, *, floor and comparisons: with contraction off, host and device produce the
// same bits, so a volume generated on the GPU equals the one the CPU oracle generates.
#pragma once
#include <math.h>
#include <stdint.h>
#include "philox.h"

namespace mrp {

MRP_HD void fbm_make_perm(uint32_t seed, int perm[512]) {
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {
        u32x4 c = {{(uint32_t)i, 0x7ea1u, 0u, 0u}};
        u32x4 r = philox4x32_10(c, seed, 0xfb3u);
        int j = (int)(r.v[0] % (uint32_t)(i + 1));
        int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    for (int i = 0; i < 256; ++i) perm[256 + i] = perm[i];
}

MRP_HD float fbm_fade(float t) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

MRP_HD float fbm_grad(int h, float x, float y, float z) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    h &= 15;
    float u = h < 8 ? x : y;
    float v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}

MRP_HD float fbm_lerp(float t, float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    return a + t * (b - a);
}

MRP_HD float fbm_noise(const int* perm, float x, float y, float z) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int X = ((int)fx) & 255, Y = ((int)fy) & 255, Z = ((int)fz) & 255;
    x -= fx; y -= fy; z -= fz;
    float u = fbm_fade(x), v = fbm_fade(y), w = fbm_fade(z);
    int A = perm[X] + Y, AA = perm[A] + Z, AB = perm[A + 1] + Z;
    int B = perm[X + 1] + Y, BA = perm[B] + Z, BB = perm[B + 1] + Z;
    return fbm_lerp(w,
                    fbm_lerp(v, fbm_lerp(u, fbm_grad(perm[AA], x, y, z), fbm_grad(perm[BA], x - 1.0f, y, z)),
                             fbm_lerp(u, fbm_grad(perm[AB], x, y - 1.0f, z), fbm_grad(perm[BB], x - 1.0f, y - 1.0f, z))),
                    fbm_lerp(v, fbm_lerp(u, fbm_grad(perm[AA + 1], x, y, z - 1.0f), fbm_grad(perm[BA + 1], x - 1.0f, y, z - 1.0f)),
                             fbm_lerp(u, fbm_grad(perm[AB + 1], x, y - 1.0f, z - 1.0f),
                                      fbm_grad(perm[BB + 1], x - 1.0f, y - 1.0f, z - 1.0f))));
}

// Raw fBm value at voxel (x,y,z) of an N^3 lattice (cell centres).
MRP_HD float fbm_voxel(const int* perm, int N, int octaves, float base_freq, int x, int y, int z) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    float px = ((float)x + 0.5f) / (float)N;
    float py = ((float)y + 0.5f) / (float)N;
    float pz = ((float)z + 0.5f) / (float)N;
    float sum = 0.0f, amp = 1.0f, freq = base_freq;
    for (int o = 0; o < octaves; ++o) {
        sum = sum + amp * fbm_noise(perm, px * freq + 0.31f * (float)o, py * freq + 0.17f * (float)o,
                                    pz * freq + 0.73f * (float)o);
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return sum;
}

// Normalised, thresholded density from a raw value and the grid's raw min/max.
MRP_HD float fbm_density(float raw, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
#endif
    float u = (raw - lo) / (hi - lo);
    return u < 0.4f ? 0.0f : (u - 0.4f) / 0.6f;
}

}  // namespace mrp
