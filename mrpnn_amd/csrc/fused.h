// Shared types of the batched predictor (march + compaction, active-list MLP kernels).
#pragma once
#include <hip/hip_runtime.h>

#include "dev_common.hpp"

namespace mrp {

// View-uniform inputs of the MFMA network kernels, computed once per view on the host with
// the device's formulas (normalize3, inv3, IEEE division) instead of once per 16-pixel tile.
struct ViewConst {
    v3 L;         // normalize(light_dir)                (render.cuh:812)
    v3 LX;        // normalize(L): LXMain                (render.cuh:694-708,823)
    v3 invLX;     // inv3(LX)                            (vector.cu:19)
    v3 sr;        // scatter_rate / 1.001                (render.cuh:843)
    v3 sp;        // sr^4                                (radiancePredict.cu:324)
    float alpha64;  // alpha / 64
    int need_y;     // channel y differs from x: a second tail evaluation
    int need_z;     // channel z differs from x and y: a third tail evaluation
};
void view_const_init(const ViewDev& vw, ViewConst* c);

// One view of a batch as the device sees it.  Its pixels are marched in tiles of 64 (one
// wavefront each) whose shape, the same for every view of a batch, the autotuner chooses.
struct ViewSlot {
    ViewDev vw;
    ViewConst vc;
    float4* out;  // width*height (rgb, distance)
    int vol;      // index into the batch's DevVolume table
    int tile0;    // first 64-pixel tile of this view in the flattened tile grid
    int tiles_x;  // tiles per row
    int checker;  // 0: every pixel; 1 + p: only pixels with (row + col + p) even, the others untouched
    int tile_lw;  // log2 of the tile width (2..6): tiles of (1 << tile_lw) x (64 >> tile_lw) pixels
};

// One active pixel after the march: 64 bytes.
struct ActiveRec {
    int view;
    int pixel;
    float pos[3];
    float X[3];
    float Y[3];
    float Z[3];
    float dist;
    int pad;
};
static_assert(sizeof(ActiveRec) == 64, "ActiveRec must be 64 bytes");

hipError_t launch_march_compact(const DevVolume* vols, const ViewSlot* slots, int n_views, int n_tiles,
                                ActiveRec* recs, unsigned* counter, unsigned long long* taps, hipStream_t s);
hipError_t launch_predict_active_fp32(const DevVolume* vols, const ViewSlot* slots, const StencilEntry* stencil,
                                      const float* W, const ActiveRec* recs, const unsigned* counter, int capacity,
                                      hipStream_t s);
hipError_t launch_reset_counters(unsigned* counter, unsigned long long* taps, hipStream_t s);

// mlp_mfma.hip: packed MFMA weights (see pack_mfma_weights) and the kernels.
struct MfmaNet;  // device pointers, defined in mlp_mfma.h
}  // namespace mrp
