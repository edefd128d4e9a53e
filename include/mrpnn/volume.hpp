// VolumeRender: the reference's public C++ class (core/volume.hpp:45-155) rebuilt on the C ABI
// of libmrpnn_hip.so (include/mrp_c_api.h).  Same names, argument order, defaults and
// meaning for the members on the predictor path; errors throw mrpnn::Error instead of
// printing and carrying on (core/volume.cu:15).
//
// Deliberate differences, each a bug or a CUDA-only mechanism of the reference:
//   * VolumeRender(int) allocates the device grid it renders from (the reference never
//     assigns datas_dev there, core/volume.cu:948-957) and no 1024^3 scratch array
//     (core/volume.cu:951-952);
//   * Update_TR(..., CPU) always bakes on the GPU;
//   * Render(size, ..., RPNN, ...) runs the RPNN network in the reference's FP32 arithmetic
//     (mrp_render_rpnn); the interactive Render(target, histo, ..., RPNN) in the interactive
//     precision (MRP_FP32, else the BF16 network);
//   * Render(..., PT, ...) returns the mean radiance with w = 0 (the reference's w is a hit
//     fraction used only by its GUI); its next-event transmittance is the cached stochastic
//     bake of Update_CachedTR (768^3) as the reference's, whether or not the volume was loaded
//     with cacheTR (the reference needs cacheTR = true there, or it writes outside its array);
//   * the random draws of GetSamples / GetTr / GetTrs / Render(PT) come from Philox streams
//     (fixed seeds, so repeated calls repeat, as the reference's per-thread cuRAND seeds do);
//   * GetHGLut reproduces the reference's host lookup, including its exchanged interpolation
//     weights (the v fraction blends along u and the u fraction along v, core/volume.cu:1438-1440).
#pragma once
#include <hip/hip_vector_types.h>

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mrp_c_api.h"

namespace mrpnn {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// core/volume.hpp:12-19: per-pixel histogram of the interactive renderer (132 bytes)
#define HISTO_SIZE MRP_HISTO_SIZE
typedef mrp_histogram Histogram;

// omp.hpp:8 (x, y, z integer coordinates; u, v, w = (i + 0.5) / resolution)
typedef std::function<float(int x, int y, int z, float u, float v, float w)> FillFunc;

// Process-wide context shared by all volumes, so they can be batched together.
mrp_ctx* default_context();

class VolumeRender {
public:
    enum RenderType { PT, RPNN, MRPNN };

    float max_density = 0.00001f;  // volume.hpp:104 -- set it to max(grid) before Update()
    float min_val = 0.0f;

    explicit VolumeRender(int resolution);
    // LoadData (core/volume.cu:970-1149): a named dataset (CLOUD0/1, MODEL0/1) from <path>.bin
    // with its fixed max density, or a .txt / .vox grid (written back as <path>.bin) / .bin file
    explicit VolumeRender(std::string path, bool cacheTR = false);
    ~VolumeRender();
    VolumeRender(const VolumeRender&) = delete;
    VolumeRender& operator=(const VolumeRender&) = delete;

    void LoadData(std::string path, bool cacheTR);
    void SetData(int x, int y, int z, float value);
    void SetDatas(FillFunc func);
    void Update(bool process_mips = true, bool need_reset_datas_dev = true);
    void Update_TR(float3 lightDir, float alpha = 64.0f, bool CPU = false);
    void UpdateHGLut(float g);

    // Environment map (core/volume.cu:1698-1736): a Radiance .hdr file; sky pixels and escaping
    // paths see it, and with it set each MRPNN frame lights the cloud either by the sun or by one
    // direction of the map (core/volume.cu:494-509, 552-568)
    void SetHDRI(std::string path);
    void SetEnvExp(float exp);

    void SetTrScale(float scale);
    void SetScatterRate(float rate);
    void SetScatterRate(float3 rate);
    // Extension: vertical field of view (degrees) of Render/Camera; 90 reproduces the reference.
    void SetFov(float degrees);

    std::vector<float4> Render(int2 size, float3 ori, float3 up, float3 right, float3 lightDir,
                               RenderType rt = RenderType::PT, float g = 0.857f, float alpha = 1,
                               float3 lightColor = {1, 1, 1}, int multiScatter = 512, int sampleNum = 1024);
    // Interactive progressive frame into caller-owned device buffers (core/volume.cu:550-628):
    // target size.x*size.y float3, histo_buffer one Histogram per pixel, target2 0xAARRGGBB.
    // Synchronises the device before returning, as the reference.
    void Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, float3 ori, float3 up,
                float3 right, float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1,
                float g = 0, int randseed = 0, RenderType rt = RenderType::PT, int toneType = 2, bool denoise = false);
    // Training-data gather (core/volume.cu:630-708): size.x * size.y rows of MRP_TILE_ROW (773)
    // floats into the caller's device buffer out_buffer, one per pixel of a size.x x size.y
    // camera, each summarising an 8 x 8 grid of sub-rays (mrp_gather_tiles); pixelOffset shifts
    // the grid by eighths of a pixel.  Updates the HG LUT and the TR mips for (g, lightDir, alpha)
    // first.  Asynchronous on the library's stream, as the reference's kernel launch; randseed is
    // unused, as in the reference.
    void GatherData(float* out_buffer, int2 size, float3 ori, float3 up, float3 right, float3 lightDir,
                    float alpha = 1, float g = 0, int2 pixelOffset = {0, 0}, int randseed = 0);
    void SetCheckboard(bool checkboard);
    void SetExposure(float exp);
    // Extension: network arithmetic of the interactive MRPNN frames (MRP_FP32 = the reference's,
    // MRP_BF16 / MRP_FP8 / MRP_BF16_1T the MFMA forms).
    void SetInteractivePrecision(int precision);

    // Host probes of the pyramids (core/volume.cu:1767-1803, 1424-1442): trilinear lookups at
    // cell centres of density mip `mip` (zero border, max(0, .)) at a position in the unit box
    // (pos) or its texture coordinate (uv); fractional mips blend levels int(mip) and +1;
    // TR mip `mip` (clamped) of the last Update_TR, points outside the box looking along the
    // light to its entry (1 when the light misses it); the HG LUT of the last UpdateHGLut at
    // (cos, cone angle in radians).  The first probe after Update / Update_TR / UpdateHGLut (or
    // a render that rebuilt them) copies the pyramid to the host.
    float DensityAtPosition(int mip, float3 pos);
    float TrAtPosition(int mip, float3 pos, float3 lightDir);
    float DensityAtPosition(float mip, float3 pos);
    float DensityAtUV(int mip, float3 uv);
    float DensityAtUV(float mip, float3 uv);
    float GetHGLut(float cos, float angle);
    // Stochastic transmittance probes (core/volume.cu:381-403, 795-830): GetTr averages
    // Tr(ori -> lightDir) * HG(dot(dir, lightDir), g) over 32 * ceil(sampleNum / 32) draws;
    // GetTrs returns lightColor * pi * the mean of sampleNum draws per (ori, dir).
    float3 GetTr(float3 ori, float3 dir, float3 lightDir, float alpha = 1, float g = 0, int sampleNum = 1) const;
    std::vector<float3> GetTrs(float alpha, std::vector<float3> ori, std::vector<float3> dir, float3 lightDir,
                               float3 lightColor, float g = 0, int sampleNum = 1) const;
    // Point-radiance training targets (core/volume.cu:710-780, core/render.cuh:615-654; the
    // caller is dataGen.cu:328): per task the radiance arriving at point ori[i] averaged over
    // sampleNum paths that start in the HG(g[i]) lobe around dir[i] (mrp_point_radiance).
    std::vector<float3> GetSamples(std::vector<float> alpha, std::vector<float3> ori, std::vector<float3> dir,
                                   std::vector<float3> lightDir, std::vector<float> g, std::vector<float> scatter,
                                   float3 lightColor = {1, 1, 1}, int multiScatter = 1, int sampleNum = 1) const;

    int resolution() const { return resolution_; }
    mrp_volume* handle() const { return vol_; }
    const std::vector<float>& datas() const { return datas_; }

private:
    // core/volume.cu:1669-1696: the path tracer's cached transmittance for (lightDir, alpha)
    void Update_CachedTR(float3 lightDir, float alpha);
    // host copies of the pyramids for the probes, with the device state they were taken from
    std::vector<float> mips_host_, tr_host_, lut_host_;
    uint32_t mips_host_gen_ = 0, tr_host_gen_ = 0;
    float tr_host_key_[4] = {0, 0, 0, -1};
    float lut_host_g_ = -100.0f;
    const float* host_mips();
    const float* host_tr();
    const float* host_lut();
    int resolution_ = 0;
    std::vector<float> datas_;
    mrp_volume* vol_ = nullptr;
    float3 scatter_rate_ = {1.001f, 1.001f, 1.001f};  // render.cuh:17
    float tr_scale_ = 1.0f;
    float tan_half_fov_ = 1.0f;
    uint32_t frame_counter_ = 0;
    mrp_interactive* interactive_ = nullptr;
    bool checkboard_ = true;  // core/volume.hpp:50
    float exposure_ = 1.0f;   // core/volume.cu:52
    int interactive_precision_ = MRP_FP32;
    int iw_ = 0, ih_ = 0;
    bool has_env_ = false;
    int rand_cpu_ = 0;           // core/volume.cu:482
    uint32_t env_draws_ = 0;     // the stream of the sun / environment choice
    float env_draw();            // uniform [0, 1) in place of rand() / RAND_MAX
    void env_light(float3& lightDir, float3& lightColor);
    void create(int resolution);
};

}  // namespace mrpnn
