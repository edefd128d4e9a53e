// Tile feature gather: the training-data path of the reference, VolumeRender::GatherData
// (core/volume.cu:630-708) -> GatherSamples (core/volume.cu:257-283) -> NNGatherSamples
// (core/render.cuh:1049-1142) -> PrepareRay (core/render.cuh:1006-1047) -> GetSampleT
// (core/render.cuh:464-486) -> PopulateInputData / SetupInputData2
// (core/radiancePredict.cu:487-565).
//
// Per pixel 64 sub-rays on an 8x8 grid around the pixel centre, grouped 2x2 into 16 groups;
// each group keeps the sub-ray hitting nearest (survival 0.75 threshold) or farthest
// (survival 0.25) in a checkerboard, and the 192 stencil points are gathered around those
// 16 samples.  One 773-float row per pixel: [X_Val 0:192 | X_Val_Hg 192:384 |
// X_Val_Sub 384:576 | gamma, d_gamma_right, d_gamma_up, horizontal t diff, vertical t diff |
// per sub-ray (mid t, t range) 581:709 | per sub-ray opacity 709:773].
//
// The reference runs all 64 optical-depth marches of a pixel serially in one thread; here a
// 64-lane wavefront is one pixel and each lane marches one sub-ray, the group choices and the
// order-dependent sums (average direction, t differences) are redone in the reference's order
// by one lane from LDS, and the 192 points are spread three per lane.
#include "../../include/mrp_c_api.h"
#include "capi_internal.h"
#include "dev_common.hpp"
#include "kernels.h"

namespace mrp {

constexpr int kTileRow = 773;

// View-uniform inputs, formed on the host with the same float operations as the oracle.
struct TileArgs {
    v3 ori, fwd, right, up;  // fwd = normalize(-ori)
    v3 right_off, up_off;    // -(right * sx) / (4 W), -(up * sy) / (4 H): one eighth-pixel step
    v3 L;                    // normalize(light_dir)
    float sx, sy, alpha, g;
    int W, H, off_x, off_y;
};

__global__ __launch_bounds__(256) void k_tile_gather(DevVolume vol, const StencilEntry* __restrict__ stencil,
                                                     TileArgs a, float* __restrict__ out) {
    __shared__ float s_ray[4][64][6];    // per sub-ray: direction, chosen t, ts.x + ts.y, ts.y > 0
    __shared__ float s_samp[4][16][4];   // the 16 group samples: direction, t
    __shared__ float s_basis[4][10];     // X, Y, Z, active
    __shared__ int s_list[4][17];        // indices of the samples that hit, then their count
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pix = blockIdx.x * 4 + wv;
    const bool ok = pix < a.W * a.H;  // no early return: every wave reaches the barriers
    const int row = ok ? pix / a.W : 0, col = ok ? pix % a.W : 0;
    float* ro = out + (size_t)(ok ? pix : 0) * kTileRow;

    // GatherSamples: pixel direction (not normalised), offset in eighths of a pixel
    const float u = 1.0f - ((float)col + (float)a.off_x / 8.0f + 0.5f) / (float)a.W;
    const float vv = 1.0f - ((float)row + (float)a.off_y / 8.0f + 0.5f) / (float)a.H;
    const v3 dir = a.fwd + (a.right * ((u * 2.0f - 1.0f) * a.sx)) + (a.up * ((vv * 2.0f - 1.0f) * a.sy));

    // PrepareRay for this lane's sub-ray: grid cell (row8, col8), group (xi, yi)
    const int row8 = lane >> 3, col8 = lane & 7;
    const int xi = col8 >> 1, yi = row8 >> 1;
    const int x = -3 + 2 * xi, y = -3 + 2 * yi;
    const float ox = (col8 & 1) ? (float)x + 0.5f : (float)x - 0.5f;
    const float oy = (row8 & 1) ? (float)y + 0.5f : (float)y - 0.5f;
    const int samp = xi + 4 * yi;
    const bool do25 = ((samp + (yi & 1)) & 1) != 0;
    const v3 sd = normalize3(dir + a.right_off * ox + a.up_off * oy);
    Ts4 ts = {0.0f, 0.0f, 0.0f, 0.0f};
    if (ok) ts = get_sample_t(vol, a.ori, sd, a.alpha);
    const float w = do25 ? ts.y : ts.x;
    const bool hit = w != 0.0f;
    {
        const float rough = 1.0f / 3.0f;
        const float min_t = ts.z * rough, max_t = fmaxf(ts.x, ts.y) * rough;
        const int gid = 16 * row8 + 2 * col8;
        if (ok) {
            ro[581 + gid] = (min_t + max_t) * 0.5f;
            ro[582 + gid] = max_t - min_t;
            ro[581 + 128 + gid / 2] = ts.w;
        }
    }
    s_ray[wv][lane][0] = sd.x;
    s_ray[wv][lane][1] = sd.y;
    s_ray[wv][lane][2] = sd.z;
    s_ray[wv][lane][3] = w;
    s_ray[wv][lane][4] = ts.x + ts.y;
    s_ray[wv][lane][5] = ts.y > 0.0f ? 1.0f : 0.0f;
    // extent of the sub-rays that hit (order-free: min / max)
    float mnx = hit ? ox : 3.402823466e38f, mny = hit ? oy : 3.402823466e38f;
    float mxx = hit ? ox : -3.402823466e38f, mxy = hit ? oy : -3.402823466e38f;
    for (int o = 32; o > 0; o >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, o));
        mny = fminf(mny, __shfl_xor(mny, o));
        mxx = fmaxf(mxx, __shfl_xor(mxx, o));
        mxy = fmaxf(mxy, __shfl_xor(mxy, o));
    }
    __syncthreads();

    if (lane < 16) {  // group choice (NNGatherSamples, render.cuh:1072-1092): lane = sampId
        const int gx = lane & 3, gy = lane >> 2;
        const int base = 16 * gy + 2 * gx;  // sub-ray lane of (x - 0.5, y - 0.5)
        const int cand[4] = {base, base + 1, base + 8, base + 9};
        const bool g25 = ((lane + (gy & 1)) & 1) != 0;
        int best = cand[0];
        for (int k = 1; k < 4; ++k) {
            const float cw = s_ray[wv][cand[k]][3], tw = s_ray[wv][best][3];
            if ((cw != 0.0f && ((cw < tw) != g25)) || tw == 0.0f) best = cand[k];
        }
        const int fill = lane / 2 + (g25 ? 8 : 0);
        for (int c = 0; c < 4; ++c) s_samp[wv][fill][c] = s_ray[wv][best][c];
    }
    if (lane == 0) {
        // average direction and t differences, summed in the reference's order: groups by
        // sampId, the four sub-rays of a group in PrepareRay call order
        v3 avg = mk3(0.0f, 0.0f, 0.0f);
        float hor = 0.0f, vert = 0.0f;
        for (int sid = 0; sid < 16; ++sid) {
            const int gx = sid & 3, gy = sid >> 2;
            const int base = 16 * gy + 2 * gx;
            const int cand[4] = {base, base + 1, base + 8, base + 9};
            float h = 0.0f;
            for (int k = 0; k < 4; ++k) {
                const float* r = s_ray[wv][cand[k]];
                if (r[3] != 0.0f) {
                    const v3 d = mk3(r[0], r[1], r[2]);
                    avg = avg + ((r[5] != 0.0f) ? d * 2.0f : d);
                    h += r[4];
                }
            }
            hor += (-3 + 2 * gx > 0) ? h : -h;
            vert += (-3 + 2 * gy > 0) ? h : -h;
        }
        const bool active = mnx != 3.402823466e38f;
        // GetMatrixFromNormal(seed, avgDir, 0) (render.cuh:694-708): with zero spread the cap
        // direction is UniformSampleSphere(0, 0) = (0, 0, 1); where the reference would loop
        // forever (|dot| < 1e-5) the basis keeps (0, 1, 0), (0, 0, 1)
        v3 X = normalize3(avg), Y = mk3(0.0f, 1.0f, 0.0f), Z = mk3(0.0f, 0.0f, 1.0f);
        const v3 r = uniform_sample_sphere(0.0f, 0.0f);
        if (!(fabsf(dot3(r, X)) <= 0.00001f)) {  // |d| < 1e-5 in double: float(1e-5) is the largest float below
            Y = normalize3(cross3(X, r));
            Z = cross3(Y, X);
        }
        s_basis[wv][0] = X.x; s_basis[wv][1] = X.y; s_basis[wv][2] = X.z;
        s_basis[wv][3] = Y.x; s_basis[wv][4] = Y.y; s_basis[wv][5] = Y.z;
        s_basis[wv][6] = Z.x; s_basis[wv][7] = Z.y; s_basis[wv][8] = Z.z;
        s_basis[wv][9] = active ? 1.0f : 0.0f;
        if (ok) {
            float* g = ro + 3 * kPC;
            if (active) {  // render.cuh:1114-1134
                const float mnx2 = mnx - 0.5f, mny2 = mny - 0.5f, mxx2 = mxx + 0.5f, mxy2 = mxy + 0.5f;
                const v3 bl = normalize3(dir + a.right_off * mnx2 + a.up_off * mny2);
                const v3 rc = normalize3(dir + a.right_off * mxx2);
                const v3 uc = normalize3(dir + a.up_off * mxy2);
                auto sc = [](float v) { return fmaxf(-1.0f, fminf(1.0f, v)); };
                g[0] = acosf(sc(dot3(X, a.L)));
                const float gbl = acosf(sc(dot3(bl, a.L)));
                g[1] = acosf(sc(dot3(rc, a.L))) - gbl;
                g[2] = acosf(sc(dot3(uc, a.L))) - gbl;
                g[3] = hor;
                g[4] = vert;
            } else {
                for (int k = 0; k < 5; ++k) g[k] = 0.0f;
            }
        }
    }
    __syncthreads();
    if (lane == 0) {  // samples that hit, in index order (SetupInputData2's round robin)
        int m = 0;
        for (int i = 0; i < 16; ++i)
            if (s_samp[wv][i][3] != 0.0f) s_list[wv][m++] = i;
        s_list[wv][16] = m;
    }
    __syncthreads();
    if (!ok) return;

    const bool active = s_basis[wv][9] != 0.0f;
    if (!active) {  // the reference writes out[0] = -1000 and leaves the row; here zeros
        for (int i = lane; i < 3 * kPC; i += 64) ro[i] = (i == 0) ? -1000.0f : 0.0f;
        return;
    }
    PixelState ps;
    ps.X = mk3(s_basis[wv][0], s_basis[wv][1], s_basis[wv][2]);
    ps.Y = mk3(s_basis[wv][3], s_basis[wv][4], s_basis[wv][5]);
    ps.Z = mk3(s_basis[wv][6], s_basis[wv][7], s_basis[wv][8]);
    ps.L = a.L;
    ps.LX = a.L;  // SetupInputData2 uses LightDir for both the LUT cosine and the shadow term
    ps.active = 1;
    ps.taps = 0;
    ps.dist = 0.0f;
    const int m = s_list[wv][16];
    for (int i = lane; i < kPC; i += 64) {
        float xv, xh, xs;
        if (i < 16) {  // radiancePredict.cu:491-509: the samples themselves
            const float* sm = s_samp[wv][i];
            if (sm[3] != 0.0f) {
                const v3 d = mk3(sm[0], sm[1], sm[2]);
                const v3 sp = a.ori + d * sm[3];
                xv = logf(1.0f + a.alpha * mip_density(vol, 0, sp) / 64.0f);
                xh = logf(henyey_greenstein(dot3(d, a.L), a.g) + 1.0f);
                xs = shadow_term_tr(vol, sp, a.L, 0);
            } else {
                xv = xh = xs = 0.0f;
            }
        } else {  // radiancePredict.cu:511-546: stencil point i around sample list[(i-16) % m]
            const float* sm = s_samp[wv][s_list[wv][(i - 16) % m]];
            ps.pos = a.ori + mk3(sm[0], sm[1], sm[2]) * sm[3];
            const v3 f = gather_point(vol, stencil[i], i, ps, a.g, a.alpha);
            xv = f.x;
            xh = f.y;
            xs = f.z;
        }
        ro[i] = xv;
        ro[kPC + i] = xh;
        ro[2 * kPC + i] = xs;
    }
}

hipError_t launch_tile_gather(const DevVolume& vol, const StencilEntry* stencil, const mrp_view* v, int off_x,
                              int off_y, float* out, hipStream_t s) {
    TileArgs a;
    a.ori = mk3(v->ori[0], v->ori[1], v->ori[2]);
    a.right = mk3(v->right[0], v->right[1], v->right[2]);
    a.up = mk3(v->up[0], v->up[1], v->up[2]);
    a.fwd = normalize3(-a.ori);
    a.W = v->width;
    a.H = v->height;
    a.sy = v->tan_half_fov;
    a.sx = v->tan_half_fov * ((float)a.W / (float)a.H);
    a.right_off = (-(a.right * a.sx)) / ((float)a.W * 4.0f);
    a.up_off = (-(a.up * a.sy)) / ((float)a.H * 4.0f);
    a.L = normalize3(mk3(v->light_dir[0], v->light_dir[1], v->light_dir[2]));
    a.alpha = v->alpha;
    a.g = v->g;
    a.off_x = off_x;
    a.off_y = off_y;
    const int n = a.W * a.H;
    hipLaunchKernelGGL(k_tile_gather, dim3((n + 3) / 4), dim3(256), 0, s, vol, stencil, a, out);
    return hipGetLastError();
}

}  // namespace mrp

using namespace mrp;

// VolumeRender::GatherData (core/volume.cu:630-708): LUT for g and TR mips for (light, alpha)
// brought up to date as the reference's UpdateHGLut / Update_TR calls do, then one row of
// 773 floats per pixel.
extern "C" int mrp_gather_tiles(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int pixel_offset_x,
                                int pixel_offset_y, float* dev_out) {
    if (!c || !v || !vw || !dev_out) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    rc = mrp_prepare_view(c, v, vw);
    if (rc) return rc;
    hipError_t e = launch_tile_gather(mrp_dev_volume(v), c->d_stencil, vw, pixel_offset_x, pixel_offset_y, dev_out,
                                      c->stream);
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "tile gather: %s", hipGetErrorString(e));
    return MRP_OK;
}
