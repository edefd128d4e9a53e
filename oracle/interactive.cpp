// Oracle of the interactive progressive path (TEST INFRASTRUCTURE ONLY): the per-pixel work of
// VolumeRender::Render(float3* target, Histogram*, unsigned* target2, ...) (core/volume.cu:550-628)
// after the frame's radiance is known.
//   RenderCamera<false> blend + histogram   core/volume.cu:109-131
//   RenderCamera<true> target write         core/volume.cu:133-150
//   Compare, UnRoll                         core/volume.cu:230-255
//   Denoise<bool>                           core/volume.cu:285-337
//   ReprojectionDenoise                     core/volume.cu:339-358
//   Gamma, ACES                             core/render.cuh:1144-1236
// The frame's radiance is an input (rgb per pixel, the renderer's raw radiance), so these
// functions can be checked on any image.
#include <math.h>
#include <stdint.h>

#include "mrp_oracle.hpp"

using namespace oracle;

namespace {

constexpr int kHisto = 10;
constexpr int kBins = 3 * kHisto;
struct Histo {
    float bin[kBins];
    float total;
    float x2;
    float x;
};
static_assert(sizeof(Histo) == 132, "Histogram layout");

bool in_set(int i, int j, int checker, int flip) { return !checker || ((((i + flip) % 2) + j) % 2) == 0; }
float saturate_ref(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
f3 lerp3(f3 a, f3 b, float v) { return mk(lerpf_(a.x, b.x, v), lerpf_(a.y, b.y, v), lerpf_(a.z, b.z, v)); }
struct M33 {
    f3 x, y, z;
};
f3 mul(const M33& m, f3 v) { return mk(dot(m.x, v), dot(m.y, v), dot(m.z, v)); }
f3 gamma22(f3 c) {
    const float e = 1.0f / 2.2f;
    return mk(powf(c.x, e), powf(c.y, e), powf(c.z, e));
}
f3 aces(f3 color) {
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
    const f3 AP1_RGB2Y = {0.272229, 0.674082, 0.0536895};
    f3 acescg = mul(AP0_2_AP1, mul(SRGB_2_AP0, color));  // unity_to_ACES, ACES_to_ACEScg
    float tmp = dot(acescg, AP1_RGB2Y);
    acescg = lerp3(mk(tmp, tmp, tmp), acescg, 0.96);
    const float a = 278.5085, b = 10.7772, c = 293.6045, d = 88.7122, e = 80.6889;
    f3 x = acescg;
    f3 rgb_post = mk((x.x * (x.x * a + b)) / (x.x * (x.x * c + d) + e), (x.y * (x.y * a + b)) / (x.y * (x.y * c + d) + e),
                     (x.z * (x.z * a + b)) / (x.z * (x.z * c + d) + e));
    // darkSurround_to_dimSurround: XYZ -> xyY, Y^0.9811, back (render.cuh:1160-1193)
    f3 XYZ = mul(AP1_2_XYZ, rgb_post);
    float divisor = (float)fmax((double)dot(XYZ, mk(1, 1, 1)), 1e-4);
    f3 xyY = mk(XYZ.x / divisor, XYZ.y / divisor, XYZ.y);
    xyY.z = (float)fmin(fmax((double)xyY.z, 0.0), 65504.0);
    xyY.z = powf(xyY.z, 0.9811f);
    float m = xyY.z / fmaxf(xyY.y, 1e-4f);
    XYZ = mk(xyY.x, xyY.z, (1.0f - xyY.x - xyY.y));
    XYZ.x *= m;
    XYZ.z *= m;
    f3 lin = mul(XYZ_2_AP1, XYZ);
    tmp = dot(lin, AP1_RGB2Y);
    lin = lerp3(mk(tmp, tmp, tmp), lin, 0.93);
    XYZ = mul(AP1_2_XYZ, lin);
    XYZ = mul(D60_2_D65, XYZ);
    lin = mul(XYZ_2_REC709, XYZ);
    return gamma22(lin);
}
uint32_t tonemap_pack(f3 res, float exposure, int tone) {
    f3 val = res * exposure;
    if (tone == 1) val = gamma22(val);
    else if (tone == 2) val = aces(val);
    const uint32_t r = (uint32_t)(255.0f * saturate_ref(val.x));
    const uint32_t g = (uint32_t)(255.0f * saturate_ref(val.y));
    const uint32_t b = (uint32_t)(255.0f * saturate_ref(val.z));
    return 0xff000000u | (r << 16) | (g << 8) | b;
}
int unroll(int r, int c, int W, int H) {  // UnRoll: x = row clamped to [0, H-1], y = column to [0, W-1]
    c = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
    r = r < 0 ? 0 : (r > H - 1 ? H - 1 : r);
    return c + W * r;
}
float compare(const Histo& x, const Histo& y) {
    float nx = x.total, ny = y.total;
    float sqrt_y_x = sqrtf(ny / nx), sqrt_x_y = sqrtf(nx / ny);
    float p = 0.0f, res = 0.0f;
    for (int i = 0; i < kBins; ++i) {
        float hx = x.bin[i], hy = y.bin[i];
        if (hx != 0.0f || hy != 0.0f) {
            p++;
            float t = sqrt_y_x * hx - sqrt_x_y * hy;
            res += t * t / (hx + hy);
        }
    }
    return p == 0.0f ? 0.0f : res / p;
}
f3 ld(const float* t, int id) { return mk(t[3 * id], t[3 * id + 1], t[3 * id + 2]); }

}  // namespace

extern "C" {

// RenderCamera<false> after the path: rad4 holds the frame's radiance (rgb, w) per pixel.
void orc_pt_accumulate(float* target, void* histo_v, const float* rad4, int W, int H, int checker, int flip,
                       int fnum) {
    Histo* histo = static_cast<Histo*>(histo_v);
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            if (!in_set(i, j, checker, flip)) continue;
            const int idx = i * W + j;
            f3 res = mk(rad4[4 * idx], rad4[4 * idx + 1], rad4[4 * idx + 2]);
            res = mk(fmaxf(0.0f, res.x), fmaxf(0.0f, res.y), fmaxf(0.0f, res.z));
            Histo& h = histo[idx];
            if (fnum == 0) h = Histo{};
            float lerp_rate = 1.0f / (1 + (fnum));
            f3 t = lerp3(ld(target, idx), res, lerp_rate);
            target[3 * idx] = t.x;
            target[3 * idx + 1] = t.y;
            target[3 * idx + 2] = t.z;
            res = mk(res.x / (res.x + 1), res.y / (res.y + 1), res.z / (res.z + 1));
            h.total += 1;
            const int bx = int(fminf(res.x, 0.999f) * kHisto), by = int(fminf(res.y, 0.999f) * kHisto),
                      bz = int(fminf(res.z, 0.999f) * kHisto);
            h.bin[bx] += 1;
            h.bin[by + kHisto] += 1;
            h.bin[bz + kHisto * 2] += 1;
            float l = dot(res, mk(1, 1, 1));
            h.x = lerpf_(h.x, l, 1.0f / (1 + fnum));
            h.x2 = lerpf_(h.x2, l * l, 1.0f / (1 + fnum));
        }
}

// RenderCamera<true>: max(0, rgb) of the frame into the target on the frame's set.
void orc_predict_commit(float* target, const float* rad4, int W, int H, int checker, int flip) {
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            if (!in_set(i, j, checker, flip)) continue;
            const int idx = i * W + j;
            target[3 * idx] = fmaxf(0.0f, rad4[4 * idx]);
            target[3 * idx + 1] = fmaxf(0.0f, rad4[4 * idx + 1]);
            target[3 * idx + 2] = fmaxf(0.0f, rad4[4 * idx + 2]);
        }
}

void orc_denoise(const float* target, const void* histo_v, uint32_t* rgba, int W, int H, int frame_num, int flip,
                 int denoise, float exposure, int tone) {
    const Histo* histo = static_cast<const Histo*>(histo_v);
#pragma omp parallel for schedule(dynamic, 16)
    for (int id = 0; id < W * H; ++id) {
        const int r = id / W, c = id % W;
        f3 res;
        if (frame_num == 0) {
            if ((((r + flip) % 2) + c) % 2 == 0) {
                res = ld(target, id);
            } else {
                res = ld(target, unroll(r - 1, c, W, H)) + ld(target, unroll(r + 1, c, W, H)) +
                      ld(target, unroll(r, c - 1, W, H)) + ld(target, unroll(r, c + 1, W, H));
                res = res / 4;
            }
        } else {
            float variance = fabsf(histo[id].x2 - histo[id].x * histo[id].x);
            if (denoise && variance > 0.01) {
                const Histo& center = histo[unroll(r, c, W, H)];
                res = mk(0, 0, 0);
                float ws = 0;
                for (int i = -5; i <= 5; i++)
                    for (int j = -5; j <= 5; j++) {
                        int t = unroll(r + i, c + j, W, H);
                        float w = fmaxf(0.0f, 1.0f - 1.2f * compare(center, histo[t]));  // single precision as the library
                        res = res + ld(target, t) * w;
                        ws += w;
                    }
                res = res / ws;
            } else {
                res = ld(target, id);
            }
        }
        rgba[id] = tonemap_pack(res, exposure, tone);
    }
}

void orc_reprojection_denoise(const float* target, void* histo_v, uint32_t* rgba, int n, float exposure, int tone) {
    Histo* histo = static_cast<Histo*>(histo_v);
    for (int id = 0; id < n; ++id) {
        histo[id].total = target[3 * id];
        histo[id].x = target[3 * id + 1];
        histo[id].x2 = target[3 * id + 2];
        rgba[id] = tonemap_pack(ld(target, id), exposure, tone);
    }
}

void orc_tonemap(const float* rgb, int n, float exposure, int tone, uint32_t* rgba, float* mapped3) {
    for (int i = 0; i < n; ++i) {
        f3 v = mk(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
        rgba[i] = tonemap_pack(v, exposure, tone);
        if (mapped3) {
            f3 m = v * exposure;
            if (tone == 1) m = gamma22(m);
            else if (tone == 2) m = aces(m);
            mapped3[3 * i] = m.x;
            mapped3[3 * i + 1] = m.y;
            mapped3[3 * i + 2] = m.z;
        }
    }
}

float orc_histo_compare(const void* x, const void* y) {
    return compare(*static_cast<const Histo*>(x), *static_cast<const Histo*>(y));
}

}  // extern "C"
