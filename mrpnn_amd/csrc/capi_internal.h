// Internal state behind the C ABI handles; shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mrp_c_api.h"
#include "dev_common.hpp"

struct FusedState;  // fused.hip

struct mrp_ctx {
    int device = 0;
    hipDeviceProp_t props{};
    hipStream_t stream = nullptr;
    std::vector<float> weights_host;
    float* d_weights = nullptr;  // fp32, header order
    mrp::StencilEntry* d_stencil = nullptr;
    float* d_lut = nullptr;  // 512x512
    bool lut_valid = false;
    float lut_g = 0.0f;
    uint32_t lut_seed = 0;
    float* d_plain_out = nullptr;
    size_t plain_out_bytes = 0;
    bool timing = false;
    bool pending_plain_timing = false;
    hipEvent_t ev[4]{};
    mrp_stats stats{};
    FusedState* fused = nullptr;
};

struct mrp_volume {
    mrp_ctx* ctx = nullptr;
    int N = 0;
    float max_density = 0.00001f;
    float* d_grid = nullptr;
    float* d_mips = nullptr;
    float* d_tr = nullptr;
    float* mip[9]{};
    float* trm[8]{};
    bool mips_valid = false;
    bool tr_valid = false;
    float tr_light[3]{};
    float tr_alpha = 0.0f;
};

int mrp_fail(int code, const char* fmt, ...);
int mrp_check_view(const mrp_view* vw);
mrp::ViewDev mrp_to_dev(const mrp_view* v);
mrp::DevVolume mrp_dev_volume(const mrp_volume* v);
int mrp_prepare_view(mrp_ctx* c, mrp_volume* v, const mrp_view* vw);
int mrp_volume_update_tr_impl(mrp_volume* v, const float L[3], float alpha);

// fused.hip
int mrp_fused_init(mrp_ctx* c);
void mrp_fused_free(mrp_ctx* c);
int mrp_fused_collect_stats(mrp_ctx* c);
int mrp_fused_predict_features(mrp_ctx* c, const float* dev_feats, int n, const float sr[3], float g, int precision,
                               float* dev_rgb);
