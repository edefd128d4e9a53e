// Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw: "Parallel random numbers:
// as easy as 1, 2, 3", SC'11), written from the published round function.  Host and device.
//
// It replaces the reference's cuRAND XORWOW streams (core/render.cuh:40-58), whose output
// cannot be reproduced outside CUDA.  Streams are addressed by (key, counter) so the same
// pixel draws the same numbers on the GPU and in the CPU oracle.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MRP_HD __host__ __device__ inline
#else
#define MRP_HD inline
#endif

namespace mrp {

struct u32x4 { uint32_t v[4]; };

MRP_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

// Ten rounds of Philox-4x32 with the published multipliers and Weyl key increments.
MRP_HD u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(M0, ctr.v[0], hi0, lo0);
        philox_mulhilo(M1, ctr.v[2], hi1, lo1);
        u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// Uniform float in the open interval (0,1): (2*(x>>9)+1) * 2^-24, i.e. the 2^23 midpoints
// of [0,1) on a 2^-23 grid.  Exact in fp32, so host and device agree bit for bit, and never
// 0 or 1 (safe for log(1-u) and log(u)).
MRP_HD float u01(uint32_t x) {
    return (float)((x >> 9) * 2u + 1u) * 5.9604644775390625e-8f;
}

// Stream purposes (counter word 1).  Counter words: {pixel, purpose, a, b}.
enum : uint32_t {
    RNG_JITTER = 1u,      // pixel jitter, a = view, b = 0
    RNG_BASIS_MAIN = 2u,  // ray basis attempts, a = view, b = attempt
    RNG_LUT = 3u,         // HG LUT per-texel jitter, a = 0, b = 0
    RNG_PT = 4u,          // path tracer: a = sample index, b = block of 4 draws
    RNG_POINT = 5u,       // point-radiance paths: pixel word = task, a = sample index
    RNG_BAKE = 6u,        // cached TR bake: pixel word = texel, a = estimate index
    RNG_TR = 7u,          // transmittance probes: pixel word = task, a = sample index
};

// Sequential draws from one (key, {pixel, purpose, a, *}) stream.
struct PhiloxStream {
    uint32_t k0, k1, c0, c1, c2, c3;
    uint32_t buf[4];
    int idx;
    MRP_HD PhiloxStream(uint32_t key0, uint32_t key1, uint32_t pixel, uint32_t purpose, uint32_t a)
        : k0(key0), k1(key1), c0(pixel), c1(purpose), c2(a), c3(0), idx(4) {}
    MRP_HD float next() {
        if (idx == 4) {
            u32x4 c = {{c0, c1, c2, c3}};
            u32x4 r = philox4x32_10(c, k0, k1);
            buf[0] = r.v[0]; buf[1] = r.v[1]; buf[2] = r.v[2]; buf[3] = r.v[3];
            ++c3;
            idx = 0;
        }
        // select chain instead of a dynamic register-array index (keeps it out of scratch)
        uint32_t x = idx == 0 ? buf[0] : (idx == 1 ? buf[1] : (idx == 2 ? buf[2] : buf[3]));
        ++idx;
        return u01(x);
    }
};

}  // namespace mrp
