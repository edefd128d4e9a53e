// The reference's README flow (README.md:72-93 / main.cu) written against mrpnn::VolumeRender
// and mrpnn::Camera: fill a grid with SetDatas, Update, render the predictor, the RPNN
// comparison network and the path tracer to PNGs.  Usage: demo <out_dir> [resolution]
#include <math.h>
#include <stdio.h>

#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "mrpnn/camera.hpp"

using namespace mrpnn;

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
        Camera cam(v, "demo", float3{1.2f, 0.6f, 1.2f}, 96);
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
        printf("active %d mean_pred %.6f mean_pt %.6f lit_pt %d lit_pred %d active_rpnn %d mean_rpnn %.6f\n", act,
               s / img.size(), sp / pt.size(), lit_pt, lit_pred, act_rp, srp / rp.size());
        return act > 0 ? 0 : 3;
    } catch (const Error& e) {
        fprintf(stderr, "mrpnn error %d: %s\n", e.code, e.what());
        return 2;
    }
}
