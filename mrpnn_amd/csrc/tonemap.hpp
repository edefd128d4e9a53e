// Tone maps of the reference's display path, for the device (interactive frames) and the host
// (Camera::Gamma / Camera::ACES): Gamma 1/2.2 and the fitted ACES curve (core/render.cuh:1144-1236,
// core/camera.cu:73-162).
#pragma once
#include "dev_common.hpp"

namespace mrp {

MRP_HD float lerpf_ref(float a, float b, float v) { return a * (1.0f - v) + b * v; }  // vector.cu:33
MRP_HD float saturatef_ref(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }            // vector.cu:23

// float3x3 * v: the dot of each row with v (vector.cu:261-263)
struct M33 {
    v3 x, y, z;
};
MRP_HD v3 mul(const M33& m, v3 v) { return mk3(dot3(m.x, v), dot3(m.y, v), dot3(m.z, v)); }
MRP_HD v3 lerp3(v3 a, v3 b, float v) {
    return mk3(lerpf_ref(a.x, b.x, v), lerpf_ref(a.y, b.y, v), lerpf_ref(a.z, b.z, v));
}
MRP_HD v3 gamma22(v3 c) {  // Gamma (render.cuh:1144-1146)
    const float e = 1.0f / 2.2f;
    return mk3(powf(c.x, e), powf(c.y, e), powf(c.z, e));
}

// ACES (render.cuh:1195-1236): sRGB -> AP0 -> AP1, 96 % saturation, the RRT+ODT rational fit,
// dark-to-dim surround (Y^0.9811 in xyY), 93 % saturation, AP1 -> XYZ, D60 -> D65 (Bradford),
// XYZ -> Rec.709, gamma 1/2.2.  The matrices are the published ACES / Rec.709 ones the
// reference lists; its double constants become float on conversion, as here.
MRP_HD v3 aces(v3 color) {
    const M33 AP1_2_XYZ = {{0.6624541811, 0.1340042065, 0.1561876870},
                           {0.2722287168, 0.6740817658, 0.0536895174},
                           {-0.0055746495, 0.0040607335, 1.0103391003}};
    const M33 SRGB_2_AP0 = {{0.4397010, 0.3829780, 0.1773350},
                            {0.0897923, 0.8134230, 0.0967616},
                            {0.0175440, 0.1115440, 0.8707040}};
    const M33 AP0_2_AP1 = {{1.4514393161, -0.2365107469, -0.2149285693},
                           {-0.0765537734, 1.1762296998, -0.0996759264},
                           {0.0083161484, -0.0060324498, 0.9977163014}};
    const M33 XYZ_2_AP1 = {{1.6410233797, -0.3248032942, -0.2364246952},
                           {-0.6636628587, 1.6153315917, 0.0167563477},
                           {0.0117218943, -0.0082844420, 0.9883948585}};
    const M33 D60_2_D65 = {{0.98722400, -0.00611327, 0.0159533},
                           {-0.00759836, 1.00186000, 0.0053302},
                           {0.00307257, -0.00509595, 1.0816800}};
    const M33 XYZ_2_REC709 = {{3.2409699419, -1.5373831776, -0.4986107603},
                              {-0.9692436363, 1.8759675015, 0.0415550574},
                              {0.0556300797, -0.2039769589, 1.0569715142}};
    const v3 AP1_RGB2Y = mk3(0.272229, 0.674082, 0.0536895);
    const v3 acescg0 = mul(AP0_2_AP1, mul(SRGB_2_AP0, color));
    float tmp = dot3(acescg0, AP1_RGB2Y);
    const v3 x = lerp3(mk3(tmp, tmp, tmp), acescg0, 0.96);
    const float a = 278.5085, b = 10.7772, c = 293.6045, d = 88.7122, e = 80.6889;
    const v3 rgb_post = mk3((x.x * (x.x * a + b)) / (x.x * (x.x * c + d) + e),
                            (x.y * (x.y * a + b)) / (x.y * (x.y * c + d) + e),
                            (x.z * (x.z * a + b)) / (x.z * (x.z * c + d) + e));
    // darkSurround_to_dimSurround (render.cuh:1172-1193)
    v3 XYZ = mul(AP1_2_XYZ, rgb_post);
    const float divisor = fmaxf(XYZ.x + XYZ.y + XYZ.z, 1e-4f);  // dot(XYZ, {1,1,1})
    float xx = XYZ.x / divisor, yy = XYZ.y / divisor, YY = XYZ.y;
    YY = fminf(fmaxf(YY, 0.0f), 65504.0f);
    YY = powf(YY, 0.9811f);
    const float m = YY / fmaxf(yy, 1e-4f);
    XYZ = mk3(xx * m, YY, (1.0f - xx - yy) * m);
    v3 lin = mul(XYZ_2_AP1, XYZ);
    tmp = dot3(lin, AP1_RGB2Y);
    lin = lerp3(mk3(tmp, tmp, tmp), lin, 0.93);
    XYZ = mul(D60_2_D65, mul(AP1_2_XYZ, lin));
    return gamma22(mul(XYZ_2_REC709, XYZ));
}

}  // namespace mrp
