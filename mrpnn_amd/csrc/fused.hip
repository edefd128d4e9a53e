// Batched predictor: many views of several resident volumes per call.
//
// Stage 1, k_march_compact: one 64-lane wavefront per 64-pixel tile (8x8 by default; 4x16,
//   16x4, 32x2 or 64x1 as chosen), the tiles of all views flattened into one grid: camera
//   ray, first-hit march, basis; inactive pixels are written (0,0,0,0) at once, active ones
//   are appended to a compact list with one atomic per wavefront, so the expensive stages see
//   dense work.
// Stage 2, gather + MLP over the active list:
//   FP32: one thread per active pixel, the reference's scalar network (parity path);
//   BF16 / FP8: the MFMA kernels of mlp_mfma.hip (16 pixels per wavefront tile).
#include "capi_internal.h"
#include "dev_common.hpp"
#include "dev_mlp_fp32.hpp"
#include "bake_order.h"
#include "fused.h"

namespace mrp {

__global__ __launch_bounds__(64) void k_march_compact(const DevVolume* __restrict__ vols, const ViewSlot* __restrict__ slots,
                                                      int n_views, ActiveRec* __restrict__ recs,
                                                      unsigned* __restrict__ counter,
                                                      unsigned long long* __restrict__ taps_total) {
    // locate the view of this tile (uniform across the block)
    const int tile = blockIdx.x;
    int v = 0;
    while (v + 1 < n_views && tile >= slots[v + 1].tile0) ++v;
    const ViewSlot& sl = slots[v];
    const ViewDev vw = sl.vw;
    const int lw = sl.tile_lw;
    // tiles in rings around the image centre, where the camera looks at the cloud's centre and
    // rays cross the most volume: the longest tiles (hundreds of dependent steps, as long as the
    // whole march) start first instead of wherever the row-major order reaches them
    const int tiles_y = (vw.height + (64 >> lw) - 1) / (64 >> lw);
    const int lt = ring_tile(tile - sl.tile0, sl.tiles_x, tiles_y, (vw.width >> 1) >> lw, (vw.height >> 1) / (64 >> lw));
    const int tx = lt % sl.tiles_x, ty = lt / sl.tiles_x;
    const int col = (tx << lw) + (threadIdx.x & ((1 << lw) - 1));
    const int row = ty * (64 >> lw) + (threadIdx.x >> lw);
    const bool valid = row < vw.height && col < vw.width && (sl.checker == 0 || ((row + col + sl.checker - 1) & 1) == 0);
    PixelState st;
    st.active = 0;
    st.taps = 0;
    if (valid) st = pixel_state(vols[sl.vol], vw, row, col);
    const int act = valid && st.active;
    const unsigned long long mask = __ballot(act);
    const int lane = threadIdx.x & 63;
    const int cnt = __popcll(mask);
    // the tile's active pixels take whole groups of 16 records, so every 16-record MFMA tile
    // belongs to one view; unused records of the last group get pixel = -1
    const int padded = (cnt + 15) & ~15;
    unsigned base = 0;
    if (lane == 0 && cnt) base = atomicAdd(counter, (unsigned)padded);
    base = __shfl(base, 0);
    const size_t pix = (size_t)row * vw.width + col;
    if (lane >= cnt && lane < padded) {
        ActiveRec d;
        d.view = v;
        d.pixel = -1;
        d.pos[0] = d.pos[1] = d.pos[2] = 0.0f;
        d.X[0] = 1.0f; d.X[1] = d.X[2] = 0.0f;
        d.Y[1] = 1.0f; d.Y[0] = d.Y[2] = 0.0f;
        d.Z[2] = 1.0f; d.Z[0] = d.Z[1] = 0.0f;
        d.dist = 0.0f;
        d.pad = 0;
        recs[base + lane] = d;
    }
    if (act) {
        const unsigned my = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
        ActiveRec r;
        r.view = v;
        r.pixel = (int)pix;
        r.pos[0] = st.pos.x; r.pos[1] = st.pos.y; r.pos[2] = st.pos.z;
        r.X[0] = st.X.x; r.X[1] = st.X.y; r.X[2] = st.X.z;
        r.Y[0] = st.Y.x; r.Y[1] = st.Y.y; r.Y[2] = st.Y.z;
        r.Z[0] = st.Z.x; r.Z[1] = st.Z.y; r.Z[2] = st.Z.z;
        r.dist = st.dist;
        r.pad = 0;
        recs[my] = r;
    } else if (valid) {  // sky: SkyBox(dir), zero without an environment map (render.cuh:829,874)
        const DevVolume& vol = vols[sl.vol];
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (vol.env) {
            const v3 sky = sky_box(vol, normalize3(pixel_ray(vw, row, col, (uint32_t)pix)));
            o = make_float4(fmaxf(0.0f, sky.x), fmaxf(0.0f, sky.y), fmaxf(0.0f, sky.z), 0.0f);
        }
        sl.out[pix] = o;
    }
    unsigned long long t = (unsigned long long)st.taps;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0 && t) atomicAdd(taps_total, t);
}

__device__ inline PixelState state_from_rec(const ActiveRec& r, const ViewDev& vw) {
    PixelState st;
    st.pos = mk3(r.pos[0], r.pos[1], r.pos[2]);
    st.X = mk3(r.X[0], r.X[1], r.X[2]);
    st.Y = mk3(r.Y[0], r.Y[1], r.Y[2]);
    st.Z = mk3(r.Z[0], r.Z[1], r.Z[2]);
    v3 light = normalize3(vw.light_dir);
    st.L = light;
    st.LX = normalize3(light);
    st.dist = r.dist;
    st.active = 1;
    st.taps = 0;
    return st;
}

// FP32 stage 2: the reference network per active pixel (same arithmetic as the plain kernel).
__global__ __launch_bounds__(64) void k_predict_active_fp32(const DevVolume* __restrict__ vols,
                                                            const ViewSlot* __restrict__ slots,
                                                            const StencilEntry* __restrict__ stencil,
                                                            const float* __restrict__ W,
                                                            const ActiveRec* __restrict__ recs,
                                                            const unsigned* __restrict__ counter) {
    const unsigned i = blockIdx.x * 64 + threadIdx.x;
    if (i >= *counter) return;
    const ActiveRec r = recs[i];
    if (r.pixel < 0) return;  // padding record
    const ViewSlot& sl = slots[r.view];
    const ViewDev vw = sl.vw;
    const DevVolume& vol = vols[sl.vol];
    PixelState st = state_from_rec(r, vw);
    float X[kPC], Hh[kPC], S[kPC];
    for (int k = 0; k < kPC; ++k) {
        v3 f = gather_point(vol, stencil[k], k, st, vw.g, vw.alpha);
        X[k] = f.x;
        Hh[k] = f.y;
        S[k] = f.z;
    }
    Fp32Mlp m;
    m.W = W;
    m.X = X;
    m.S = S;
    m.H = Hh;
    v3 sr = vw.scatter_rate / 1.001f;
    v3 rgb = m.run(gamma_of(st), vw.g, sr);
    v3 o = (rgb * vw.light_color) * vw.scatter_rate;
    sl.out[r.pixel] = make_float4(fmaxf(0.0f, o.x), fmaxf(0.0f, o.y), fmaxf(0.0f, o.z), fmaxf(0.0f, r.dist));
}

// Counters are cleared by a kernel (not hipMemsetAsync) so the sequence can be captured in a
// HIP graph and replayed.
__global__ void k_reset_counters(unsigned* counter, unsigned long long* taps) {
    if (threadIdx.x == 0) {
        *counter = 0u;
        *taps = 0ull;
    }
}
hipError_t launch_reset_counters(unsigned* counter, unsigned long long* taps, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counter, taps);
    return hipGetLastError();
}

hipError_t launch_march_compact(const DevVolume* vols, const ViewSlot* slots, int n_views, int n_tiles,
                                ActiveRec* recs, unsigned* counter, unsigned long long* taps, hipStream_t s) {
    hipLaunchKernelGGL(k_march_compact, dim3(n_tiles), dim3(64), 0, s, vols, slots, n_views, recs, counter, taps);
    return hipGetLastError();
}
hipError_t launch_predict_active_fp32(const DevVolume* vols, const ViewSlot* slots, const StencilEntry* stencil,
                                      const float* W, const ActiveRec* recs, const unsigned* counter, int capacity,
                                      hipStream_t s) {
    hipLaunchKernelGGL(k_predict_active_fp32, dim3((capacity + 63) / 64), dim3(64), 0, s, vols, slots, stencil, W,
                       recs, counter);
    return hipGetLastError();
}

}  // namespace mrp
