// The reference's README flow (README.md:72-93 / main.cu) written against mrpnn::VolumeRender
// and mrpnn::Camera: fill a grid with SetDatas, Update, render the predictor, the RPNN
// comparison network and the path tracer to PNGs, gather training rows into a device buffer as
// main.cu:30-38 does, and check the class results against the C ABI calls they wrap.
// Usage: demo <out_dir> [resolution]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "mrpnn/camera.hpp"

using namespace mrpnn;

// The camera's basis (protected in the reference's Camera, kept so here) for the C ABI views.
struct BasisCamera : Camera {
    using Camera::Camera;
    void basis(float o[3], float u[3], float r[3]) const {
        const float3 a[3] = {ori, up, right};
        float* d[3] = {o, u, r};
        for (int i = 0; i < 3; ++i) {
            d[i][0] = a[i].x;
            d[i][1] = a[i].y;
            d[i][2] = a[i].z;
        }
    }
};

static mrp_view capi_view(const BasisCamera& cam, int w, int h, float3 sun, float alpha, float g, float sr) {
    mrp_view v;
    memset(&v, 0, sizeof v);
    cam.basis(v.ori, v.up, v.right);
    v.light_dir[0] = sun.x;
    v.light_dir[1] = sun.y;
    v.light_dir[2] = sun.z;
    v.light_color[0] = v.light_color[1] = v.light_color[2] = 1.0f;
    v.scatter_rate[0] = v.scatter_rate[1] = v.scatter_rate[2] = sr;
    v.alpha = alpha;
    v.g = g;
    v.tan_half_fov = 1.0f;
    v.width = w;
    v.height = h;
    return v;
}

int main(int argc, char** argv) {
    std::string dir = argc > 1 ? argv[1] : ".";
    int res = argc > 2 ? atoi(argv[2]) : 64;
    try {
        VolumeRender v(res);
        v.SetDatas([](int, int, int, float u, float vv, float w) {
            float dx = u - 0.5f, dy = vv - 0.5f, dz = w - 0.5f;
            float r = sqrtf(dx * dx + dy * dy + dz * dz);
            return r < 0.35f ? 1.0f - r / 0.35f : 0.0f;  // soft ball
        });
        v.max_density = 1.0f;
        v.Update();
        v.SetScatterRate(0.99f);
        BasisCamera cam(v, "demo", float3{1.2f, 0.6f, 1.2f}, 96);
        float3 sun = {0.5f, 0.8f, 0.33f};
        float l = sqrtf(sun.x * sun.x + sun.y * sun.y + sun.z * sun.z);
        sun = float3{sun.x / l, sun.y / l, sun.z / l};
        auto img = cam.RenderToFile(dir + "/demo_mrpnn", sun, float3{1, 1, 1}, 32.0f, 512, 0.857f, 1, Camera::None,
                                    VolumeRender::MRPNN);
        double s = 0;
        int act = 0;
        for (auto& p : img) {
            s += p.x;
            act += p.w > 0;
        }
        // the comparison network of the paper, in the reference's FP32 arithmetic
        auto rp = cam.RenderToFile(dir + "/demo_rpnn", sun, float3{1, 1, 1}, 32.0f, 1, 0.857f, 1, Camera::None,
                                   VolumeRender::RPNN);
        double srp = 0;
        int act_rp = 0;
        for (auto& p : rp) {
            srp += p.x;
            act_rp += p.w > 0;
        }
        auto pt = cam.RenderToFile(dir + "/demo_pt", sun, float3{1, 1, 1}, 32.0f, 64, 0.857f, 2, Camera::None,
                                   VolumeRender::PT);
        double sp = 0;
        for (auto& p : pt) sp += p.x;
        // the GUI loop (core/GUI.hpp:1020-1027): progressive frames into device buffers
        const int W = 64;
        float3* accum = nullptr;
        Histogram* histo = nullptr;
        unsigned int* rgba = nullptr;
        if (hipMalloc((void**)&accum, W * W * sizeof(float3)) != hipSuccess ||
            hipMalloc((void**)&histo, W * W * sizeof(Histogram)) != hipSuccess ||
            hipMalloc((void**)&rgba, W * W * sizeof(unsigned int)) != hipSuccess)
            return 4;
        if (hipMemset(accum, 0, W * W * sizeof(float3)) != hipSuccess) return 4;
        v.SetCheckboard(false);
        v.SetExposure(1.0f);
        int frame = 0;
        for (; frame < 3; ++frame)
            cam.Render(accum, histo, rgba, int2{W, W}, frame, sun, float3{1, 1, 1}, 32.0f, 64, 0.857f, 2,
                       VolumeRender::PT, true);
        std::vector<unsigned int> px(W * W);
        if (hipMemcpy(px.data(), rgba, px.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return 4;
        int lit_pt = 0;
        for (unsigned int p : px) lit_pt += (p & 0xffffffu) != 0 && (p >> 24) == 0xffu;
        cam.Render(accum, histo, rgba, int2{W, W}, frame, sun, float3{1, 1, 1}, 32.0f, 1, 0.857f, 2,
                   VolumeRender::MRPNN, true);
        if (hipMemcpy(px.data(), rgba, px.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return 4;
        int lit_pred = 0;
        for (unsigned int p : px) lit_pred += (p & 0xffffffu) != 0 && (p >> 24) == 0xffu;
        if (hipFree(accum) != hipSuccess || hipFree(histo) != hipSuccess || hipFree(rgba) != hipSuccess) return 4;
        // the same MRPNN frame through the C ABI the class wraps (first Render call: frame 0,
        // fresh jitter stream, seed 0): RenderToFile must return it bit for bit
        mrp_view pv = capi_view(cam, 96, 96, sun, 32.0f, 0.857f, 0.99f);
        pv.jitter = -1.0f;
        std::vector<float4> capi(img.size());
        if (mrp_render_plain(default_context(), v.handle(), &pv, 1, (float*)capi.data()) != MRP_OK) return 5;
        const int render_equal = memcmp(capi.data(), img.data(), img.size() * sizeof(float4)) == 0;
        // main.cu:30-38: training rows of a resolution/8 camera into a device buffer
        const int T = cam.resolution / 8;
        const size_t row_bytes = (192 * 3 + 5 + 64 * 3) * sizeof(float);
        float* d_buffer = nullptr;
        float* d_ref = nullptr;
        if (hipMalloc((void**)&d_buffer, T * T * row_bytes) != hipSuccess ||
            hipMalloc((void**)&d_ref, T * T * row_bytes) != hipSuccess)
            return 4;
        int gather_equal = 1, gather_active = 0;
        const int offsets[2][2] = {{0, 0}, {3, 5}};
        for (const auto& o : offsets) {
            cam.GatherData(d_buffer, int2{T, T}, sun, 32.0f, 0.857f, int2{o[0], o[1]});
            mrp_view gv = capi_view(cam, T, T, sun, 32.0f, 0.857f, 0.99f);
            gv.jitter = 0.5f;
            if (mrp_gather_tiles(default_context(), v.handle(), &gv, o[0], o[1], d_ref) != MRP_OK) return 5;
            std::vector<float> a(T * T * row_bytes / 4), b(a.size());
            if (hipDeviceSynchronize() != hipSuccess ||
                hipMemcpy(a.data(), d_buffer, T * T * row_bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(b.data(), d_ref, T * T * row_bytes, hipMemcpyDeviceToHost) != hipSuccess)
                return 4;
            gather_equal &= memcmp(a.data(), b.data(), a.size() * 4) == 0;
            for (int r = 0; r < T * T; ++r) gather_active += a[(size_t)r * 773] != -1000.0f;
        }
        if (hipFree(d_buffer) != hipSuccess || hipFree(d_ref) != hipSuccess) return 4;
        // the tone maps on a few colours, for comparison with the display path
        const float3 cs[3] = {{0.5f, 0.2f, 0.1f}, {2.0f, 1.0f, 0.25f}, {0.01f, 0.03f, 0.02f}};
        printf("tones");
        for (const float3& c : cs) {
            const float3 ga = Camera::Gamma(c), ac = Camera::ACES(c);
            printf(" %.9g %.9g %.9g %.9g %.9g %.9g", ga.x, ga.y, ga.z, ac.x, ac.y, ac.z);
        }
        printf("\n");
        // the per-frame Render of the camera: name + "MRPNN.png" with the ACES tone map
        cam.name = dir + "/demo_cam";
        cam.Render(sun, float3{1, 1, 1}, 32.0f, 512, 0.857f, 1, Camera::ACES, VolumeRender::MRPNN);
        // host probes (dataGen.cu:74,173) on the pyramids of the last MRPNN frame (TR for sun,
        // alpha 32; LUT for g 0.857), with the arrays they read written out for the test
        {
            std::vector<float> mips(19173961), tr(2396745), lut(512 * 512);
            if (mrp_volume_download_mips(v.handle(), mips.data()) != MRP_OK ||
                mrp_volume_download_tr(v.handle(), tr.data()) != MRP_OK ||
                mrp_ctx_download_hg_lut(default_context(), lut.data()) != MRP_OK)
                return 5;
            FILE* f = fopen((dir + "/probe_arrays.bin").c_str(), "wb");
            if (!f) return 6;
            fwrite(mips.data(), 4, mips.size(), f);
            fwrite(tr.data(), 4, tr.size(), f);
            fwrite(lut.data(), 4, lut.size(), f);
            fclose(f);
            const float3 pts[5] = {{0.1f, -0.05f, 0.2f}, {-0.3f, 0.12f, 0.07f}, {0.0f, 0.0f, 0.0f},
                                   {0.45f, 0.3f, -0.2f}, {0.7f, -0.6f, 0.1f}};
            printf("probes");
            for (const float3& q : pts) {
                printf(" %.9g %.9g %.9g %.9g %.9g %.9g", v.DensityAtPosition(0, q), v.DensityAtPosition(2, q),
                       v.DensityAtPosition(1.25f, q), v.DensityAtUV(1, float3{q.x + 0.5f, q.y + 0.5f, q.z + 0.5f}),
                       v.TrAtPosition(0, q, sun), v.TrAtPosition(3, q, sun));
            }
            const float hg[4][2] = {{0.3f, 0.2f}, {-0.7f, 0.9f}, {0.99f, 0.01f}, {-1.0f, 1.2f}};
            for (const auto& c : hg) printf(" %.9g", v.GetHGLut(c[0], c[1]));
            printf("\n");
            // stochastic probes: GetSamples / GetTrs equal the C ABI calls they wrap (seed 0)
            std::vector<float3> so = {{0.05f, 0.1f, -0.1f}, {-0.2f, 0.0f, 0.15f}, {0.2f, -0.25f, 0.0f}};
            std::vector<float3> sd = {{1, 0, 0}, {0, 0, 1}, {0.6f, 0.8f, 0.0f}};
            std::vector<float3> sl = {sun, sun, sun};
            std::vector<float> sa = {32, 32, 16}, sg = {0.857f, 0.5f, 0.857f}, ss = {0.99f, 0.99f, 0.9f};
            auto gs = v.GetSamples(sa, so, sd, sl, sg, ss, float3{1, 1, 1}, 64, 256);
            auto gt = v.GetTrs(32.0f, so, sd, sun, float3{1, 0.5f, 0.25f}, 0.857f, 64);
            float3 t1 = v.GetTr(so[0], sd[0], sun, 32.0f, 0.857f, 40);
            float *d_in = nullptr, *d_out = nullptr, *d_t = nullptr;
            if (hipMalloc((void**)&d_in, 3 * 15 * sizeof(float)) != hipSuccess ||
                hipMalloc((void**)&d_out, 9 * sizeof(float)) != hipSuccess ||
                hipMalloc((void**)&d_t, 3 * sizeof(float)) != hipSuccess)
                return 4;
            std::vector<float> hin;
            for (auto* arr : {&so, &sd, &sl})
                for (const float3& q : *arr) hin.insert(hin.end(), {q.x, q.y, q.z});
            hin.insert(hin.end(), sa.begin(), sa.end());
            hin.insert(hin.end(), sg.begin(), sg.end());
            hin.insert(hin.end(), ss.begin(), ss.end());
            if (hipMemcpy(d_in, hin.data(), hin.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return 4;
            const float lc[3] = {1, 1, 1}, L[3] = {sun.x, sun.y, sun.z};
            if (mrp_point_radiance(default_context(), v.handle(), 3, d_in, d_in + 9, d_in + 18, d_in + 27, d_in + 30,
                                   d_in + 33, lc, 64, 256, 0u, 1.0f, d_out, nullptr) != MRP_OK ||
                mrp_tr_samples(default_context(), v.handle(), 3, d_in, d_in + 9, L, 32.0f, 0.857f, 64, 0u, 1.0f, d_t) !=
                    MRP_OK)
                return 5;
            float hout[9], ht[3];
            if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(hout, d_out, sizeof hout, hipMemcpyDeviceToHost) ||
                hipMemcpy(ht, d_t, sizeof ht, hipMemcpyDeviceToHost) != hipSuccess)
                return 4;
            int samples_equal = memcmp(hout, gs.data(), sizeof hout) == 0;
            for (int i = 0; i < 3; ++i)
                samples_equal &= gt[i].x == 1.0f * 3.1415926535f * ht[i] && gt[i].z == 0.25f * 3.1415926535f * ht[i];
            (void)hipFree(d_in);
            (void)hipFree(d_out);
            (void)hipFree(d_t);
            printf("stochastic %.9g %.9g %.9g %.9g samples_equal %d\n", gs[0].x, gs[1].y, gs[2].z, t1.x, samples_equal);
        }
        printf("active %d mean_pred %.6f mean_pt %.6f lit_pt %d lit_pred %d active_rpnn %d mean_rpnn %.6f "
               "render_equal %d gather_equal %d gather_active %d\n", act, s / img.size(), sp / pt.size(), lit_pt,
               lit_pred, act_rp, srp / rp.size(), render_equal, gather_equal, gather_active);
        return act > 0 && render_equal && gather_equal ? 0 : 3;
    } catch (const Error& e) {
        fprintf(stderr, "mrpnn error %d: %s\n", e.code, e.what());
        return 2;
    }
}
