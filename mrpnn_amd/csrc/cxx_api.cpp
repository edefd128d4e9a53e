// mrpnn::VolumeRender / mrpnn::Camera over the C ABI (see include/mrpnn/*.hpp).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include <fstream>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mrpnn/camera.hpp"
#include "../../include/mrpnn/volume.hpp"
#include "tonemap.hpp"

namespace mrpnn {

static void check(int rc) {
    if (rc != MRP_OK) throw Error(rc, mrp_last_error());
}

mrp_ctx* default_context() {
    static std::mutex mu;
    static mrp_ctx* ctx = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!ctx) check(mrp_ctx_create(nullptr, &ctx));
    return ctx;
}

void VolumeRender::create(int resolution) {
    resolution_ = resolution;
    datas_.assign((size_t)resolution * resolution * resolution, 0.0f);
    check(mrp_volume_create(default_context(), resolution, &vol_));
}

VolumeRender::VolumeRender(int resolution) { create(resolution); }

VolumeRender::VolumeRender(std::string path, bool cacheTR) { LoadData(path, cacheTR); }

VolumeRender::~VolumeRender() {
    if (interactive_) mrp_interactive_destroy(interactive_);
    if (vol_) mrp_volume_destroy(vol_);
}

// LoadData (core/volume.cu:970-1149) through mrp_volume_read: the named datasets' fixed max
// densities, .txt / .vox parsing with the <path>.bin write-back, plain .bin files.
void VolumeRender::LoadData(std::string path, bool /*cacheTR*/) {
    float* grid = nullptr;
    int n = 0;
    float md = 0.0f;
    check(mrp_volume_read(path.c_str(), &grid, &n, &md));
    if (vol_) {
        mrp_volume_destroy(vol_);
        vol_ = nullptr;
    }
    try {
        create(n);
    } catch (...) {
        mrp_host_free(grid);
        throw;
    }
    memcpy(datas_.data(), grid, datas_.size() * sizeof(float));
    mrp_host_free(grid);
    max_density = md;
    Update();
}

void VolumeRender::SetHDRI(std::string path) {
    check(mrp_ctx_load_hdri(default_context(), path.c_str()));
    has_env_ = true;
}

void VolumeRender::SetEnvExp(float exp) { check(mrp_ctx_set_env_exp(default_context(), exp)); }

float VolumeRender::env_draw() {  // splitmix32-style hash of a counter: a uniform in [0, 1)
    uint32_t x = 0x9e3779b9u * (++env_draws_);
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}

void VolumeRender::env_light(float3& lightDir, float3& lightColor) {
    if (!has_env_) return;
    const float d[3] = {lightDir.x, lightDir.y, lightDir.z}, c[3] = {lightColor.x, lightColor.y, lightColor.z};
    float od[3], oc[3];
    int rc = mrp_ctx_env_light(default_context(), d, c, rand_cpu_, env_draw(), od, oc);
    if (rc < 0) throw Error(MRP_ERR_INVALID, "environment light");
    lightDir = float3{od[0], od[1], od[2]};
    lightColor = float3{oc[0], oc[1], oc[2]};
}

void VolumeRender::SetData(int x, int y, int z, float value) {
    datas_[((size_t)x * resolution_ + y) * resolution_ + z] = value;  // volume.cu:1178-1180
}

void VolumeRender::SetDatas(FillFunc func) {  // ParallelFill (omp.cpp:4-18), threads over x slabs
    const int r = resolution_;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            for (int i = (int)t; i < r; i += (int)nt)
                for (int j = 0; j < r; ++j)
                    for (int k = 0; k < r; ++k) {
                        float u = (float)((i + 0.5) / r), v = (float)((j + 0.5) / r), w = (float)((k + 0.5) / r);
                        datas_[((size_t)i * r + j) * r + k] = func(i, j, k, u, v, w);
                    }
        });
    for (auto& th : pool) th.join();
}

void VolumeRender::Update(bool process_mips, bool need_reset_datas_dev) {
    if (need_reset_datas_dev) check(mrp_volume_set_data(vol_, datas_.data(), max_density));
    if (process_mips) check(mrp_volume_update(vol_));
}

void VolumeRender::Update_TR(float3 lightDir, float alpha, bool /*CPU*/) {
    const float L[3] = {lightDir.x, lightDir.y, lightDir.z};
    check(mrp_volume_update_tr(vol_, L, alpha));
}

void VolumeRender::UpdateHGLut(float g) { check(mrp_ctx_update_hg_lut(default_context(), g, 1u)); }

void VolumeRender::SetTrScale(float scale) { tr_scale_ = scale; }
void VolumeRender::SetScatterRate(float rate) { scatter_rate_ = float3{rate, rate, rate}; }
void VolumeRender::SetScatterRate(float3 rate) { scatter_rate_ = rate; }
void VolumeRender::SetFov(float degrees) {
    if (!(degrees > 0.0f && degrees < 180.0f)) throw Error(MRP_ERR_INVALID, "fov must be in (0, 180)");
    tan_half_fov_ = degrees == 90.0f ? 1.0f : (float)tan(degrees * M_PI / 360.0);
}

std::vector<float4> VolumeRender::Render(int2 size, float3 ori, float3 up, float3 right, float3 lightDir,
                                         RenderType rt, float g, float alpha, float3 lightColor, int multiScatter,
                                         int sampleNum) {
    mrp_view v;
    memset(&v, 0, sizeof v);
    const float3 o[6] = {ori, up, right, lightDir, lightColor, scatter_rate_};
    float* dst[6] = {v.ori, v.up, v.right, v.light_dir, v.light_color, v.scatter_rate};
    for (int i = 0; i < 6; ++i) {
        dst[i][0] = o[i].x;
        dst[i][1] = o[i].y;
        dst[i][2] = o[i].z;
    }
    v.alpha = alpha;
    v.g = g;
    v.tan_half_fov = tan_half_fov_;
    v.jitter = -1.0f;  // a fresh sub-pixel jitter every frame, as the reference's Rand()
    v.width = size.x;
    v.height = size.y;
    v.seed = 0;
    v.frame = frame_counter_;
    frame_counter_ += (uint32_t)(sampleNum > 0 ? sampleNum : 1);
    std::vector<float4> out((size_t)size.x * size.y);
    if (rt == RenderType::MRPNN || rt == RenderType::RPNN) {
        if (rt == RenderType::MRPNN && !has_env_) {
            check(mrp_render_plain(default_context(), vol_, &v, sampleNum, (float*)out.data()));
            return out;
        }
        // frame by frame (RPNN, or an environment map: every frame draws its own light,
        // core/volume.cu:494-509; Roberts2 uses the same counter through one Render call, as
        // the reference), averaged progressively; RPNN in the reference's FP32 arithmetic
        const int frames = sampleNum > 0 ? sampleNum : 1;
        float* dev = nullptr;
        if (hipMalloc((void**)&dev, out.size() * sizeof(float4)) != hipSuccess)
            throw Error(MRP_ERR_NOMEM, "frame allocation failed");
        std::vector<float4> acc(out.size());
        int rc = MRP_OK;
        for (int f = 0; f < frames && rc == MRP_OK; ++f) {
            mrp_view vf = v;
            vf.frame = v.frame + (uint32_t)f;
            float3 ld = lightDir, lc = lightColor;
            env_light(ld, lc);
            vf.light_dir[0] = ld.x;
            vf.light_dir[1] = ld.y;
            vf.light_dir[2] = ld.z;
            vf.light_color[0] = lc.x;
            vf.light_color[1] = lc.y;
            vf.light_color[2] = lc.z;
            rc = rt == RenderType::RPNN ? mrp_render_rpnn(default_context(), vol_, &vf, MRP_FP32, dev)
                                        : mrp_render_plain_device(default_context(), vol_, &vf, dev);
            if (rc == MRP_OK && hipMemcpy(out.data(), dev, out.size() * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
                rc = MRP_ERR_HIP;
            const float rate = 1.0f / (1.0f + (float)f);  // the progressive average (volume.cu:226-227)
            for (size_t i = 0; rc == MRP_OK && i < out.size(); ++i) {
                float4& a = acc[i];
                const float4& b = out[i];
                a = f == 0 ? b
                           : float4{a.x * (1.0f - rate) + b.x * rate, a.y * (1.0f - rate) + b.y * rate,
                                    a.z * (1.0f - rate) + b.z * rate, a.w * (1.0f - rate) + b.w * rate};
            }
        }
        (void)hipFree(dev);
        check(rc);
        return acc;
    }
    // PT: sampleNum frames of 32 paths (core/volume.cu:222), NEE through the cached
    // stochastic transmittance bake (core/volume.cu:515-518, core/render.cuh:589)
    Update_CachedTR(lightDir, alpha);
    const size_t n = (size_t)size.x * size.y;
    double *mean = nullptr, *m2 = nullptr;
    if (hipMalloc((void**)&mean, n * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&m2, n * 3 * sizeof(double)) != hipSuccess) {
        hipFree(mean);
        throw Error(MRP_ERR_NOMEM, "PT accumulator allocation failed");
    }
    const int spp = 32 * (sampleNum > 0 ? sampleNum : 1);
    int rc = MRP_OK;
    for (int s0 = 0; s0 < spp && rc == MRP_OK; s0 += 32)
        rc = mrp_pt_render(default_context(), vol_, &v, s0, 32, multiScatter, 2, tr_scale_, mean, m2);
    std::vector<double> h(n * 3);
    if (rc == MRP_OK && hipMemcpy(h.data(), mean, n * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = MRP_ERR_HIP;
    hipFree(mean);
    hipFree(m2);
    check(rc);
    for (size_t i = 0; i < n; ++i) out[i] = float4{(float)h[3 * i], (float)h[3 * i + 1], (float)h[3 * i + 2], 0.0f};
    return out;
}

void VolumeRender::Update_CachedTR(float3 lightDir, float alpha) {
    const float L[3] = {lightDir.x, lightDir.y, lightDir.z};
    check(mrp_volume_bake_cached_tr(vol_, L, alpha, tr_scale_, 0u, 0));  // cached on the exact key
}

// ------------------------------------------------------------------------------------------
// Host probes (core/volume.cu:835-880, 1424-1442, 1767-1803) on host copies of the pyramids
// ------------------------------------------------------------------------------------------
namespace {
using mrp::v3;
float fetch_border_h(const float* d, int res, int x, int y, int z) {  // Sample(data, res, x, y, z)
    if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) return 0.0f;
    return fmaxf(0.0f, d[((size_t)x * res + y) * res + z]);
}
float fetch_clamp_h(const float* d, int res, int x, int y, int z) {  // SampleClamp(data, res, x, y, z)
    x = x < 0 ? 0 : (x > res - 1 ? res - 1 : x);
    y = y < 0 ? 0 : (y > res - 1 ? res - 1 : y);
    z = z < 0 ? 0 : (z > res - 1 ? res - 1 : z);
    return fmaxf(0.0f, d[((size_t)x * res + y) * res + z]);
}
// Sample / SampleClamp (data, res, uv): trilinear at cell centres, x the slowest axis
template <float (*F)(const float*, int, int, int, int)>
float trilinear_h(const float* d, int res, v3 uv) {
    const v3 p = uv * (float)res - 0.5f;
    const int x = (int)floorf(p.x), y = (int)floorf(p.y), z = (int)floorf(p.z);
    const v3 w = p - mrp::mk3((float)x, (float)y, (float)z);
    using mrp::lerp_ref;
    return lerp_ref(lerp_ref(lerp_ref(F(d, res, x, y, z), F(d, res, x, y, z + 1), w.z),
                             lerp_ref(F(d, res, x, y + 1, z), F(d, res, x, y + 1, z + 1), w.z), w.y),
                    lerp_ref(lerp_ref(F(d, res, x + 1, y, z), F(d, res, x + 1, y, z + 1), w.z),
                             lerp_ref(F(d, res, x + 1, y + 1, z), F(d, res, x + 1, y + 1, z + 1), w.z), w.y),
                    w.x);
}
size_t level_offset(int top, int m) {  // levels top^3, (top/2)^3, ... concatenated
    size_t o = 0;
    for (int k = 0; k < m; ++k) o += (size_t)(top >> k) * (top >> k) * (top >> k);
    return o;
}
v3 v3_of(float3 a) { return mrp::mk3(a.x, a.y, a.z); }

// Device buffers of the probes, freed on scope exit.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) throw Error(MRP_ERR_NOMEM, "probe allocation failed");
    }
    ~DevBuf() { (void)hipFree(p); }
    template <class T>
    T* as() const { return (T*)p; }
};
void upload(void* dst, const void* src, size_t bytes) {
    if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) throw Error(MRP_ERR_HIP, "probe upload failed");
}
}  // namespace

const float* VolumeRender::host_mips() {
    mrp_volume_info in;
    check(mrp_volume_get_info(vol_, &in));
    if (!in.mips_valid) throw Error(MRP_ERR_STATE, "Update() must run before the density probes");
    if (mips_host_.empty() || mips_host_gen_ != in.mips_generation) {
        mips_host_.resize(level_offset(256, 9));
        check(mrp_volume_download_mips(vol_, mips_host_.data()));
        mips_host_gen_ = in.mips_generation;
    }
    return mips_host_.data();
}
const float* VolumeRender::host_tr() {
    mrp_volume_info in;
    check(mrp_volume_get_info(vol_, &in));
    if (!in.tr_valid) throw Error(MRP_ERR_STATE, "Update_TR() must run before the transmittance probes");
    const float key[4] = {in.tr_light_dir[0], in.tr_light_dir[1], in.tr_light_dir[2], in.tr_alpha};
    if (tr_host_.empty() || memcmp(key, tr_host_key_, sizeof key) != 0 || tr_host_gen_ != in.mips_generation) {
        tr_host_.resize(level_offset(128, 8));
        check(mrp_volume_download_tr(vol_, tr_host_.data()));
        memcpy(tr_host_key_, key, sizeof key);
        tr_host_gen_ = in.mips_generation;
    }
    return tr_host_.data();
}
const float* VolumeRender::host_lut() {
    float g = 0.0f;
    check(mrp_ctx_get_hg_lut_g(default_context(), &g));
    if (lut_host_.empty() || g != lut_host_g_) {
        lut_host_.resize(512 * 512);
        check(mrp_ctx_download_hg_lut(default_context(), lut_host_.data()));
        lut_host_g_ = g;
    }
    return lut_host_.data();
}

float VolumeRender::DensityAtPosition(int mip, float3 pos) {  // core/volume.cu:1767-1769
    if (mip < 0 || mip > 8) throw Error(MRP_ERR_INVALID, "density mip in [0, 8]");
    return trilinear_h<fetch_border_h>(host_mips() + level_offset(256, mip), 256 >> mip, v3_of(pos) + 0.5f);
}
float VolumeRender::DensityAtUV(int mip, float3 uv) {  // core/volume.cu:1792-1794
    if (mip < 0 || mip > 8) throw Error(MRP_ERR_INVALID, "density mip in [0, 8]");
    return trilinear_h<fetch_border_h>(host_mips() + level_offset(256, mip), 256 >> mip, v3_of(uv));
}
float VolumeRender::DensityAtPosition(float mip, float3 pos) {  // core/volume.cu:1786-1790
    if (!(mip >= 0.0f && mip <= 8.0f)) throw Error(MRP_ERR_INVALID, "density mip in [0, 8]");
    const int a = (int)mip;
    const float w = mip - (float)a;
    return mrp::lerp_ref(DensityAtPosition(a, pos), a < 8 ? DensityAtPosition(a + 1, pos) : 0.0f, w);
}
float VolumeRender::DensityAtUV(float mip, float3 uv) {  // core/volume.cu:1796-1800
    if (!(mip >= 0.0f && mip <= 8.0f)) throw Error(MRP_ERR_INVALID, "density mip in [0, 8]");
    const int a = (int)mip;
    const float w = mip - (float)a;
    return mrp::lerp_ref(DensityAtUV(a, uv), a < 8 ? DensityAtUV(a + 1, uv) : 0.0f, w);
}
float VolumeRender::TrAtPosition(int mip, float3 pos, float3 lightDir) {  // core/volume.cu:1770-1784
    if (mip < 0 || mip > 7) throw Error(MRP_ERR_INVALID, "TR mip in [0, 7]");
    const float* tr = host_tr() + level_offset(128, mip);
    const int res = 128 >> mip;
    const v3 p = v3_of(pos), L = v3_of(lightDir);
    if (p.x < -0.5f || p.y < -0.5f || p.z < -0.5f || p.x > 0.5f || p.y > 0.5f || p.z > 0.5f) {
        const float off = mrp::ray_box_offset(p, L);
        if (off >= 0.0f) return trilinear_h<fetch_clamp_h>(tr, res, p + 0.5f + L * off);
        return 1.0f;  // TR_MUL
    }
    return trilinear_h<fetch_clamp_h>(tr, res, p + 0.5f);
}
float VolumeRender::GetHGLut(float cs, float angle) {  // core/volume.cu:1424-1442, weights as there
    const float* lut = host_lut();
    const float u = (float)(cs * 0.5 + 0.5);
    const float v = (float)(angle / (3.1415926535 * 60.0 / 180.0));
    auto cl = [](int i) { return i < 0 ? 0 : (i > 511 ? 511 : i); };
    const int u0 = cl((int)(u * 512)), v0 = cl((int)(v * 512));
    const int u1 = cl((int)(u * 512) + 1), v1 = cl((int)(v * 512) + 1);
    const float a00 = lut[u0 + v0 * 512], a10 = lut[u1 + v0 * 512];
    const float a01 = lut[u0 + v1 * 512], a11 = lut[u1 + v1 * 512];
    const float fv = v * 512 - floorf(v * 512), fu = u * 512 - floorf(u * 512);
    const float ax0 = mrp::lerp_ref(a00, a10, fv), ax1 = mrp::lerp_ref(a01, a11, fv);
    return mrp::lerp_ref(ax0, ax1, fu);
}

// ------------------------------------------------------------------------------------------
// Stochastic probes and point-radiance targets on the GPU
// ------------------------------------------------------------------------------------------
float3 VolumeRender::GetTr(float3 ori, float3 dir, float3 lightDir, float alpha, float g, int sampleNum) const {
    const int draws = 32 * (((sampleNum > 0 ? sampleNum : 1) + 31) / 32);  // whole 32-thread groups
    const float od[6] = {ori.x, ori.y, ori.z, dir.x, dir.y, dir.z};
    DevBuf d_od(sizeof od), d_out(sizeof(float));
    upload(d_od.p, od, sizeof od);
    const float L[3] = {lightDir.x, lightDir.y, lightDir.z};
    check(mrp_tr_samples(default_context(), vol_, 1, d_od.as<float>(), d_od.as<float>() + 3, L, alpha, g, draws, 0u,
                         tr_scale_, d_out.as<float>()));
    float r = 0.0f;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&r, d_out.p, sizeof r, hipMemcpyDeviceToHost) != hipSuccess)
        throw Error(MRP_ERR_HIP, "GetTr download failed");
    return float3{r, r, r};
}

std::vector<float3> VolumeRender::GetTrs(float alpha, std::vector<float3> ori, std::vector<float3> dir, float3 lightDir,
                                         float3 lightColor, float g, int sampleNum) const {
    if (ori.size() != dir.size()) throw Error(MRP_ERR_INVALID, "GetTrs: ori and dir differ in length");
    const int n = (int)ori.size();
    std::vector<float3> res(n);
    if (n == 0) return res;
    DevBuf d_o(n * sizeof(float3)), d_d(n * sizeof(float3)), d_out(n * sizeof(float));
    upload(d_o.p, ori.data(), n * sizeof(float3));
    upload(d_d.p, dir.data(), n * sizeof(float3));
    const float L[3] = {lightDir.x, lightDir.y, lightDir.z};
    check(mrp_tr_samples(default_context(), vol_, n, d_o.as<float>(), d_d.as<float>(), L, alpha, g,
                         sampleNum > 0 ? sampleNum : 1, 0u, tr_scale_, d_out.as<float>()));
    std::vector<float> t(n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(t.data(), d_out.p, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw Error(MRP_ERR_HIP, "GetTrs download failed");
    const float k = 3.1415926535f;  // ShadowTerm_TRs: lightColor * pi * mean (core/render.cuh:679)
    for (int i = 0; i < n; ++i) res[i] = float3{lightColor.x * k * t[i], lightColor.y * k * t[i], lightColor.z * k * t[i]};
    return res;
}

std::vector<float3> VolumeRender::GetSamples(std::vector<float> alpha, std::vector<float3> ori, std::vector<float3> dir,
                                             std::vector<float3> lightDir, std::vector<float> g,
                                             std::vector<float> scatter, float3 lightColor, int multiScatter,
                                             int sampleNum) const {
    const size_t n = ori.size();
    if (dir.size() != n || lightDir.size() != n || alpha.size() != n || g.size() != n || scatter.size() != n)
        throw Error(MRP_ERR_INVALID, "GetSamples: task arrays differ in length");
    std::vector<float3> res(n);
    if (n == 0) return res;
    DevBuf d_o(n * sizeof(float3)), d_d(n * sizeof(float3)), d_l(n * sizeof(float3)), d_a(n * sizeof(float)),
        d_g(n * sizeof(float)), d_s(n * sizeof(float)), d_out(n * sizeof(float3));
    upload(d_o.p, ori.data(), n * sizeof(float3));
    upload(d_d.p, dir.data(), n * sizeof(float3));
    upload(d_l.p, lightDir.data(), n * sizeof(float3));
    upload(d_a.p, alpha.data(), n * sizeof(float));
    upload(d_g.p, g.data(), n * sizeof(float));
    upload(d_s.p, scatter.data(), n * sizeof(float));
    const float lc[3] = {lightColor.x, lightColor.y, lightColor.z};
    check(mrp_point_radiance(default_context(), vol_, (int)n, d_o.as<float>(), d_d.as<float>(), d_l.as<float>(),
                             d_a.as<float>(), d_g.as<float>(), d_s.as<float>(), lc, multiScatter > 0 ? multiScatter : 1,
                             sampleNum > 0 ? sampleNum : 1, 0u, tr_scale_, d_out.as<float>(), nullptr));
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(res.data(), d_out.p, n * sizeof(float3), hipMemcpyDeviceToHost) != hipSuccess)
        throw Error(MRP_ERR_HIP, "GetSamples download failed");
    return res;
}

void VolumeRender::SetCheckboard(bool checkboard) { checkboard_ = checkboard; }
void VolumeRender::SetExposure(float exp) { exposure_ = exp; }
void VolumeRender::SetInteractivePrecision(int precision) {
    if (precision < MRP_FP32 || precision > MRP_BF16_1T) throw Error(MRP_ERR_INVALID, "unknown precision");
    interactive_precision_ = precision;
}

void VolumeRender::Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, float3 ori,
                          float3 up, float3 right, float3 lightDir, float3 lightColor, float alpha, int multiScatter,
                          float g, int randseed, RenderType rt, int toneType, bool denoise) {
    // the renderer keeps its frame state (flip counter, last mode) across calls of one size
    if (!interactive_ || size.x != iw_ || size.y != ih_) {
        if (interactive_) mrp_interactive_destroy(interactive_);
        interactive_ = nullptr;
        check(mrp_interactive_create(default_context(), size.x, size.y, &interactive_));
        iw_ = size.x;
        ih_ = size.y;
    }
    check(mrp_interactive_set_checkerboard(interactive_, checkboard_ ? 1 : 0));
    check(mrp_interactive_set_exposure(interactive_, exposure_));
    mrp_view v;
    memset(&v, 0, sizeof v);
    const float3 o[6] = {ori, up, right, lightDir, lightColor, scatter_rate_};
    float* dst[6] = {v.ori, v.up, v.right, v.light_dir, v.light_color, v.scatter_rate};
    for (int i = 0; i < 6; ++i) {
        dst[i][0] = o[i].x;
        dst[i][1] = o[i].y;
        dst[i][2] = o[i].z;
    }
    if (rt != RenderType::PT && has_env_) {
        float3 ld = lightDir, lc = lightColor;
        env_light(ld, lc);
        v.light_dir[0] = ld.x;
        v.light_dir[1] = ld.y;
        v.light_dir[2] = ld.z;
        v.light_color[0] = lc.x;
        v.light_color[1] = lc.y;
        v.light_color[2] = lc.z;
    }
    ++rand_cpu_;  // core/volume.cu:579
    v.alpha = alpha;
    v.g = g;
    v.tan_half_fov = tan_half_fov_;
    v.jitter = -1.0f;
    v.width = size.x;
    v.height = size.y;
    // RPNN frames run its FP32 network, or its BF16 network for any MFMA precision
    const int prec = rt == RenderType::RPNN && interactive_precision_ != MRP_FP32 ? MRP_BF16 : interactive_precision_;
    check(mrp_render_interactive(interactive_, vol_, &v, rt == RenderType::PT ? MRP_RT_PT : rt == RenderType::RPNN ? MRP_RT_RPNN : MRP_RT_MRPNN,
                                 prec, multiScatter, randseed, toneType, denoise ? 1 : 0,
                                 reinterpret_cast<float*>(target), histo_buffer, target2));
    if (hipDeviceSynchronize() != hipSuccess) throw Error(MRP_ERR_HIP, "interactive frame failed");
}

void VolumeRender::GatherData(float* out_buffer, int2 size, float3 ori, float3 up, float3 right, float3 lightDir,
                              float alpha, float g, int2 pixelOffset, int /*randseed*/) {
    if (!out_buffer) throw Error(MRP_ERR_INVALID, "GatherData: null output buffer");
    mrp_view v;
    memset(&v, 0, sizeof v);
    const float3 o[4] = {ori, up, right, lightDir};
    float* dst[4] = {v.ori, v.up, v.right, v.light_dir};
    for (int i = 0; i < 4; ++i) {
        dst[i][0] = o[i].x;
        dst[i][1] = o[i].y;
        dst[i][2] = o[i].z;
    }
    v.light_color[0] = v.light_color[1] = v.light_color[2] = 1.0f;
    v.scatter_rate[0] = scatter_rate_.x;
    v.scatter_rate[1] = scatter_rate_.y;
    v.scatter_rate[2] = scatter_rate_.z;
    v.alpha = alpha;
    v.g = g;
    v.tan_half_fov = tan_half_fov_;
    v.jitter = 0.5f;
    v.width = size.x;
    v.height = size.y;
    check(mrp_gather_tiles(default_context(), vol_, &v, pixelOffset.x, pixelOffset.y, out_buffer));
}

// ------------------------------------------------------------------------------------------
static float3 from_v3(mrp::v3 a) { return float3{a.x, a.y, a.z}; }
float3 Camera::Gamma(float3 c) { return from_v3(mrp::gamma22(mrp::mk3(c.x, c.y, c.z))); }
float3 Camera::ACES(float3 c) { return from_v3(mrp::aces(mrp::mk3(c.x, c.y, c.z))); }

Camera::Camera(VolumeRender& vol, std::string nm, float3 position, int res) {
    volume = &vol;
    name = nm;
    resolution = res;
    SetPosition(position);
}

static float3 n3(float3 v) {
    float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return float3{v.x / l, v.y / l, v.z / l};
}
static float3 c3(float3 a, float3 b) {
    return float3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

void Camera::SetPosition(float3 position) {  // camera.cu:176-192
    ori = position;
    forward = n3(float3{-ori.x, -ori.y, -ori.z});
    right = n3(c3(float3{0, 1, 0}, forward));
    up = c3(forward, right);
    if (position.x == 0 && position.z == 0) {
        right = float3{1, 0, 0};
        up = position.y > 0 ? float3{0, 0, 1} : float3{0, 0, -1};
    }
}

void Camera::SetVolume(VolumeRender& vol) { volume = &vol; }

void Camera::Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, int frame,
                    float3 lightDir, float3 lightColor, float alpha, int multiScatter, float g, int tone,
                    VolumeRender::RenderType rt, bool denoise) {  // camera.cu:259-261
    volume->Render(target, histo_buffer, target2, size, ori, up, right, lightDir, lightColor, alpha, multiScatter, g,
                   frame, rt, tone, denoise);
}

void Camera::GatherData(float* out_buffer, int2 size, float3 lightDir, float alpha, float g, int2 pixelOffset) {
    volume->GatherData(out_buffer, size, ori, up, right, lightDir, alpha, g, pixelOffset);  // camera.cu:264-266
}

std::vector<float4> Camera::Render(float3 lightDir, float3 lightColor, float alpha, int multiScatter, float g,
                                   int sampleNum, ToneMapFunction tone, VolumeRender::RenderType rt, float exp) {
    std::vector<float4> img = volume->Render(int2{resolution, resolution}, ori, up, right, lightDir, rt, g, alpha,
                                             lightColor, multiScatter, sampleNum);
    std::vector<unsigned char> px((size_t)resolution * resolution * 4);
    auto sat = [](float v) { return (unsigned char)((v < 0 ? 0 : (v > 1 ? 1 : v)) * 255); };
    for (size_t i = 0; i < img.size(); ++i) {
        const float3 t = tone(float3{img[i].x * exp, img[i].y * exp, img[i].z * exp});
        px[4 * i] = sat(t.x);
        px[4 * i + 1] = sat(t.y);
        px[4 * i + 2] = sat(t.z);
        px[4 * i + 3] = 255;
    }
    const char* affix = rt == VolumeRender::RenderType::PT ? "_PT" : (rt == VolumeRender::RenderType::RPNN ? "_RPNN" : "MRPNN");
    if (!write_png_rgba(name + affix + ".png", resolution, resolution, px.data()))
        throw Error(MRP_ERR_IO, "cannot write " + name + affix + ".png");
    return img;
}

std::vector<float4> Camera::RenderToFile(std::string path, float3 lightDir, float3 lightColor, float alpha,
                                         int multiScatter, float g, int sampleNum, ToneMapFunction tone,
                                         VolumeRender::RenderType rt, float exp) {
    (void)tone;
    (void)exp;  // the reference ignores both in RenderToFile (camera.cu:243)
    std::vector<float4> img = volume->Render(int2{resolution, resolution}, ori, up, right, lightDir, rt, g, alpha,
                                             lightColor, multiScatter, sampleNum);
    std::vector<unsigned char> px((size_t)resolution * resolution * 4);
    auto sat = [](float v) { return (unsigned char)((v < 0 ? 0 : (v > 1 ? 1 : v)) * 255); };
    for (size_t i = 0; i < img.size(); ++i) {
        px[4 * i] = sat(img[i].x);
        px[4 * i + 1] = sat(img[i].y);
        px[4 * i + 2] = sat(img[i].z);
        px[4 * i + 3] = 255;
    }
    if (!write_png_rgba(path + ".png", resolution, resolution, px.data()))
        throw Error(MRP_ERR_IO, "cannot write " + path + ".png");
    return img;
}

// ------------------------------------------------------------------------------------------
// PNG: signature, IHDR, one IDAT holding a zlib stream of stored (uncompressed) deflate
// blocks, IEND.  CRC-32 (ISO 3309 polynomial) and Adler-32 as in RFC 1950/2083.
static uint32_t crc_table[256];
static void crc_init() {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}
static uint32_t crc(const unsigned char* b, size_t n, uint32_t c = 0xffffffffu) {
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ b[i]) & 0xff] ^ (c >> 8);
    return c;
}
static void be32(std::vector<unsigned char>& o, uint32_t v) {
    o.push_back(v >> 24);
    o.push_back(v >> 16);
    o.push_back(v >> 8);
    o.push_back(v);
}
static void chunk(std::vector<unsigned char>& out, const char* type, const std::vector<unsigned char>& data) {
    be32(out, (uint32_t)data.size());
    std::vector<unsigned char> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    be32(out, crc(td.data(), td.size()) ^ 0xffffffffu);
}

bool write_png_rgba(const std::string& path, int w, int h, const unsigned char* rgba) {
    static std::once_flag once;
    std::call_once(once, crc_init);
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (4 * w + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgba + (size_t)y * w * 4, rgba + (size_t)(y + 1) * w * 4);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        size_t len = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
        z.push_back(pos + len == raw.size() ? 1 : 0);
        z.push_back(len & 0xff);
        z.push_back(len >> 8);
        z.push_back(~len & 0xff);
        z.push_back((~len >> 8) & 0xff);
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    be32(z, (b << 16) | a);
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    be32(ihdr, (uint32_t)w);
    be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    return fclose(f) == 0 && ok;
}

}  // namespace mrpnn
