// Host-side launchers of the HIP kernels (one translation unit per kernel family).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mrp_c_api.h"

#include "dev_common.hpp"
#include "dev_gather_fast.hpp"

namespace mrp {

// volume_kernels.hip
hipError_t launch_fbm(float* out, int N, int octaves, float base_freq, const int* perm, uint32_t* minmax,
                      hipStream_t s);
hipError_t launch_box2(const float* src, float* dst, int res, hipStream_t s);
hipError_t launch_resample(const float* src, int sres, float* dst, int res, hipStream_t s);
hipError_t launch_tr_bake(const float* const mips[9], const float* pmips, const float4* qmips, float* const tr[8],
                          float* tr_pad,
                          const float alpha_lvl[8], v3 light, hipStream_t s);
hipError_t launch_pad_mips(const float* const mips[9], float* pmips, hipStream_t s);
hipError_t launch_quad_mips(const float* pmips, float4* qmips, hipStream_t s);
hipError_t launch_occupancy(const float* grid, int N, uint8_t* occ, int occ_n, hipStream_t s);
hipError_t launch_fast_stencil(const StencilEntry* st, const float* lut, void* fast, float* lut_rows, hipStream_t s);
hipError_t launch_hg_lut(float* lut, float g, uint32_t seed, hipStream_t s);

// predict_plain.hip
hipError_t launch_predict_plain(const DevVolume& vol, const StencilEntry* stencil, const float* W, const ViewDev& vw,
                                float* out4, float lerp_rate, int first_frame, hipStream_t s);
hipError_t launch_gather_features(const DevVolume& vol, const StencilEntry* stencil, const ViewDev& vw, float* feats,
                                  int* active, float* pos3, int* taps, hipStream_t s);
hipError_t launch_predict_features_fp32(const float* feats, int n, const float* W, v3 sr, float g, float* rgb,
                                        hipStream_t s);

// tile_gather.hip
hipError_t launch_tile_gather(const DevVolume& vol, const StencilEntry* stencil, const mrp_view* v, int off_x,
                              int off_y, float* out, hipStream_t s);

}  // namespace mrp
