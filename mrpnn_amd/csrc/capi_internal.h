// Internal state behind the C ABI handles; shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mrp_c_api.h"
#include "dev_common.hpp"

struct FusedState;  // fused.hip

// Packed pyramid sizes (floats): density mips 0..8 ((256>>m)^3), TR mips 0..7 ((128>>m)^3).
constexpr size_t kMipFloats = 19173961;
constexpr size_t kTrFloats = 2396745;

// Device timing of renders without a host sync per render: each timed render records three
// events (start, after the march, end) into a ring slot; a slot is read back when it is
// reused or when the statistics are requested.
struct TimingRing {
    static constexpr int kSlots = 32;
    hipEvent_t ev[kSlots][4]{};  // start, march done, network start, end
    bool pending[kSlots]{};
    int head = 0;
    int64_t count = 0;
    double sum_march = 0.0, sum_predict = 0.0;
    float last_march = 0.0f, last_predict = 0.0f;
};
int mrp_timing_begin(struct mrp_ctx* c, hipEvent_t** ev);  // ev = the four events of a fresh slot
int mrp_timing_collect(struct mrp_ctx* c);                  // reads back every pending slot

// One transmittance-mip set of the batched path, keyed by (volume, light direction, alpha):
// views of one volume under different suns each get their own set.
struct TrSlot {
    const struct mrp_volume* vol = nullptr;
    float light[3]{};
    float alpha = 0.0f;
    float* d_tr = nullptr;   // TR mips 0..7 (kTrFloats)
    float* d_ptr = nullptr;  // (y, z) quads of the zero-padded TR mips 0..6 (kPadTrAlloc floats)
    uint64_t stamp = 0;      // last prepare call that used it
    bool valid = false;
    hipEvent_t last_read = nullptr;  // recorded after the last batch that read this set
    bool read_pending = false;
};

struct RpnnState;

struct mrp_ctx {
    int device = 0;
    hipDeviceProp_t props{};
    hipStream_t stream = nullptr;
    std::vector<float> weights_host;
    float* d_weights = nullptr;  // fp32, header order
    mrp::StencilEntry* d_stencil = nullptr;
    float* d_lut = nullptr;  // 512x512
    void* d_fast = nullptr;  // FastStencil[192], rebuilt with the LUT
    float* d_lut_rows = nullptr;  // per stencil point the LUT row at its v (kLutRow floats each)
    bool lut_valid = false;
    float lut_g = 0.0f;
    uint32_t lut_seed = 0;
    float* d_plain_out = nullptr;
    size_t plain_out_bytes = 0;
    bool timing = false;
    bool exact_gather = false;  // fused MFMA kernels gather with the exact (plain) arithmetic
    hipEvent_t ev[4]{};
    TimingRing ring;
    mrp_stats stats{};
    FusedState* fused = nullptr;
    std::vector<TrSlot> tr_slots;  // TR cache of the batched path
    int tr_slot_cap = 64;
    uint64_t tr_stamp = 0;
    // environment map (SetHDRI / SetEnvExp): device texels, and the host copy the light
    // sampling reads
    float4* d_env = nullptr;
    std::vector<float> env_host;
    int env_w = 0, env_h = 0;
    float env_exp = 1.0f;
    // RPNN comparison network: weights and its working buffers, loaded on first use
    struct RpnnState* rpnn = nullptr;
    int rpnn_chunk = 65536;  // pixels per RPNN descriptor buffer (BF16 path)
    // LoadData's DataCache: host grids of the four named datasets, read once per context
    std::vector<float> data_cache[4];
    int data_cache_n[4] = {0, 0, 0, 0};
};

struct mrp_volume {
    mrp_ctx* ctx = nullptr;
    int N = 0;
    float max_density = 0.00001f;
    float* d_grid = nullptr;
    float* d_mips = nullptr;
    float* d_tr = nullptr;
    float* d_pmips = nullptr;  // zero-padded density mips 0..7
    float4* d_qmips = nullptr; // (y, z) quads of padded density mips 1..7 (TR bake)
    float* d_ptr = nullptr;    // (y, z) quads of the zero-padded TR mips 0..6
    uint8_t* d_occ = nullptr;  // brick emptiness for the first-hit march, occ_n^3
    int occ_n = 0;
    float* mip[9]{};
    float* trm[8]{};
    bool mips_valid = false;
    uint32_t mips_gen = 0;            // successful pyramid rebuilds
    hipEvent_t mips_ready = nullptr;  // recorded after the last pyramid rebuild
    bool tr_valid = false;
    float tr_light[3]{};
    float tr_alpha = 0.0f;
    // cached stochastic transmittance bake of the path tracer's NEE (Update_CachedTR)
    float* d_ctr = nullptr;
    int ctr_res = 0;
    bool ctr_valid = false;
    float ctr_light[3]{};
    float ctr_alpha = 0.0f, ctr_scale = 0.0f;
    uint32_t ctr_seed = 0;
};

int mrp_fail(int code, const char* fmt, ...);
int mrp_check_view(const mrp_view* vw);
mrp::ViewDev mrp_to_dev(const mrp_view* v);
mrp::DevVolume mrp_dev_volume(const mrp_volume* v);
int mrp_prepare_view(mrp_ctx* c, mrp_volume* v, const mrp_view* vw);
int mrp_volume_update_tr_impl(mrp_volume* v, const float L[3], float alpha);
int mrp_tr_bake_into(mrp_volume* v, const float L[3], float alpha, float* d_tr, float* d_ptr, hipStream_t s);
void mrp_tr_slots_drop(mrp_ctx* c, const mrp_volume* v);  // invalidate a volume's cached sets
int mrp_tr_slot_find(const mrp_ctx* c, const mrp_volume* v, const float L[3], float alpha);

// rpnn.hip
void mrp_rpnn_free(mrp_ctx* c);
// mrp_render_rpnn on one checkerboard half (checker as mrp_render_batch_checker: 0 = all
// pixels, else the pixels with (row + col + checker - 1) even); the others are not written
int mrp_render_rpnn_checker(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int precision, float* dev_out4,
                            int checker);

// fused.hip
int mrp_fused_init(mrp_ctx* c);
// mrp_render_batch with a checkerboard per view (ViewSlot::checker; nullptr = every pixel)
int mrp_render_batch_checker(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                             int precision, float* const* dev_out4, const int* checker);
void mrp_fused_free(mrp_ctx* c);
int mrp_fused_collect_stats(mrp_ctx* c);
int mrp_fused_join_aux(mrp_ctx* c);
// resolves the no-alloc scratch array on the current device (once per device)
void mrp_noalloc_setup(int device);
int mrp_fused_predict_features(mrp_ctx* c, const float* dev_feats, int n, const float sr[3], float g, int precision,
                               float* dev_rgb);
