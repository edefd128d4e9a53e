// Camera: the reference's pinhole camera (core/camera.hpp:10-43, core/camera.cu:166-257) on
// top of mrpnn::VolumeRender.  RenderToFile writes `path + ".png"` (RGBA8, saturate x 255).
#pragma once
#include <functional>
#include <string>

#include "volume.hpp"

namespace mrpnn {

typedef std::function<float3(float3 radiance)> ToneMapFunction;

class Camera {
protected:
    float3 ori;
    float3 forward;
    float3 right;
    float3 up;
    VolumeRender* volume;

public:
    // Tone maps (core/camera.cu:69-162): pow(c, 1/2.2), and the fitted ACES curve followed by it
    static float3 Gamma(float3 color);
    static float3 ACES(float3 color);
    static float3 None(float3 color) { return color; }

    int resolution;
    std::string name;

    Camera(VolumeRender& volume, std::string name = "test camera", float3 position = {0.75f, 0.75f, 0.75f},
           int resolution = 512);
    void SetPosition(float3 position);
    float3 GetPosition() const { return ori; }
    void SetVolume(VolumeRender& volume);
    // Renders the camera's square view and writes name + "_PT" / "_RPNN" / "MRPNN" + ".png" with
    // tone(radiance * exp), saturated (the body of core/camera.cu:197-226, which the reference
    // keeps commented out); returns the untoned image (rgb, distance).
    std::vector<float4> Render(float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1,
                               float g = 0, int sampleNum = 1, ToneMapFunction tone = ACES,
                               VolumeRender::RenderType rt = VolumeRender::RenderType::PT, float exp = 1.0f);
    // Returns the rendered image (rgb, distance) and writes path.png (saturated radiance: tone
    // and exp are ignored, as in core/camera.cu:229-257).
    std::vector<float4> RenderToFile(std::string path, float3 lightDir, float3 lightColor = {1, 1, 1},
                                     float alpha = 1, int multiScatter = 1, float g = 0, int sampleNum = 1,
                                     ToneMapFunction tone = ACES,
                                     VolumeRender::RenderType rt = VolumeRender::RenderType::PT, float exp = 1.0f);
    // Interactive frame of the GUI (core/camera.cu:259-261): VolumeRender's interactive Render.
    void Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, int frame,
                float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1, float g = 0,
                int tone = 2, VolumeRender::RenderType rt = VolumeRender::RenderType::PT, bool denoise = false);
    // Training rows of this camera (core/camera.cu:264-266): VolumeRender::GatherData.
    void GatherData(float* out_buffer, int2 size, float3 lightDir, float alpha = 1, float g = 0,
                    int2 pixelOffset = {0, 0});
};

// Minimal PNG writer (stored deflate blocks), RGBA8, row-major.
bool write_png_rgba(const std::string& path, int width, int height, const unsigned char* rgba);

}  // namespace mrpnn
