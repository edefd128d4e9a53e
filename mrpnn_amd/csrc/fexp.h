// exp(x) in single precision from basic IEEE operations only, for the FP32 networks of the
// device and of the CPU restatement: the same source compiled for both gives the same bits
// (every step is a correctly rounded float operation: multiply, fused multiply-add,
// round-to-integer, exact power-of-two scaling), so an FP32 network run on the GPU and on the
// host agree exactly without evaluating anything in double precision.
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions", 1980): argument
// reduction x = k ln2 + r with k = rint(x / ln2) and ln2 split into a 16-bit head (k * head is
// exact for |k| < 2^8) and a tail, so |r| <= ln2 / 2; e^r from its Taylor series to degree 7
// (truncation below 6e-9 relative for |r| <= 0.347) in Horner form; e^x = 2^k e^r.
// Accuracy: within 2 ulp of the correctly rounded result on [-87, 88] (tested against a
// double-precision exp over 4 million points).
#pragma once
#include <math.h>

#include "philox.h"  // MRP_HD

namespace mrp {

MRP_HD float exp_f32(float x) {
    if (x > 88.72283f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    const float k = rintf(x * 1.44269502f);           // x / ln2 to the nearest integer
    float r = fmaf(k, -0.693145751953125f, x);        // ln2 head: 0x1.62e400p-1, 16 significant bits
    r = fmaf(k, -1.42860677e-06f, r);                 // ln2 tail: ln2 - head, rounded to float
    // e^r = 1 + r + r^2/2! + ... + r^7/7!
    float p = 1.98412698e-04f;                        // 1/7!
    p = fmaf(p, r, 1.38888889e-03f);                  // 1/6!
    p = fmaf(p, r, 8.33333333e-03f);                  // 1/5!
    p = fmaf(p, r, 4.16666667e-02f);                  // 1/4!
    p = fmaf(p, r, 1.66666667e-01f);                  // 1/3!
    p = fmaf(p, r, 0.5f);                             // 1/2!
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

}  // namespace mrp
