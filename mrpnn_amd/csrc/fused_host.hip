// Host side of the batched predictor: view tables, active-list buffers, MFMA weight
// packing, dispatch, statistics.
#include <math.h>
#include <string.h>

#include <vector>

#include "capi_internal.h"
#include "fused.h"
#include "mlp_mfma.h"

using namespace mrp;

struct FusedState {
    ActiveRec* d_recs = nullptr;
    size_t cap = 0;  // records
    unsigned* d_counter = nullptr;
    unsigned long long* d_taps = nullptr;
    DevVolume* d_vols = nullptr;
    ViewSlot* d_slots = nullptr;
    int vols_cap = 0, slots_cap = 0;
    DevVolume* h_vols = nullptr;  // pinned staging
    ViewSlot* h_slots = nullptr;
    hipEvent_t staged = nullptr;  // staging buffers free once this completes
    bool have_counts = false;
    MfmaNet net;  // packed BF16 / FP8 weights
    int mfma_waves_per_cu = 8;
};

static int grow(void** p, size_t bytes) {
    hipFree(*p);
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) return mrp_fail(MRP_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
    return MRP_OK;
}

int mrp_fused_init(mrp_ctx* c) {
    FusedState* f = new FusedState();
    c->fused = f;
    if (hipMalloc(&f->d_counter, 16) != hipSuccess || hipMalloc(&f->d_taps, 16) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "counter allocation failed");
    if (hipEventCreateWithFlags(&f->staged, hipEventDisableTiming) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "event creation failed");
    int rc = mfma_net_create(c->weights_host.data(), &f->net);
    if (rc) return rc;
    return MRP_OK;
}

void mrp_fused_free(mrp_ctx* c) {
    FusedState* f = c->fused;
    if (!f) return;
    hipFree(f->d_recs);
    hipFree(f->d_counter);
    hipFree(f->d_taps);
    hipFree(f->d_vols);
    hipFree(f->d_slots);
    hipHostFree(f->h_vols);
    hipHostFree(f->h_slots);
    if (f->staged) hipEventDestroy(f->staged);
    mfma_net_destroy(&f->net);
    delete f;
    c->fused = nullptr;
}

int mrp_fused_collect_stats(mrp_ctx* c) {
    FusedState* f = c->fused;
    int rc = mrp_timing_collect(c);
    if (rc) return rc;
    if (f->have_counts) {
        unsigned cnt = 0;
        unsigned long long taps = 0;
        if (hipMemcpyAsync(&cnt, f->d_counter, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(&taps, f->d_taps, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return mrp_fail(MRP_ERR_HIP, "reading counters failed");
        c->stats.active_pixels = cnt;
        c->stats.march_taps = (int64_t)taps;
        f->have_counts = false;
    }
    return MRP_OK;
}

extern "C" int mrp_prepare_views(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views) {
    if (!c || !vols || !views || n_vols < 1 || n_views < 1) return mrp_fail(MRP_ERR_INVALID, "bad batch arguments");
    std::vector<int> light_of(n_vols, -1);
    for (int i = 0; i < n_views; ++i) {
        int rc = mrp_check_view(&views[i]);
        if (rc) return rc;
        int vi = views[i].volume_index;
        if (vi < 0 || vi >= n_vols || !vols[vi]) return mrp_fail(MRP_ERR_INVALID, "view %d: bad volume_index", i);
        if (views[i].g != views[0].g) return mrp_fail(MRP_ERR_INVALID, "all views of a batch must share g");
        if (light_of[vi] >= 0) {
            const mrp_view& a = views[light_of[vi]];
            if (memcmp(a.light_dir, views[i].light_dir, sizeof a.light_dir) != 0 || a.alpha != views[i].alpha)
                return mrp_fail(MRP_ERR_INVALID, "views %d and %d share a volume but not (light, alpha)",
                                light_of[vi], i);
            continue;
        }
        light_of[vi] = i;
        rc = mrp_prepare_view(c, vols[vi], &views[i]);
        if (rc) return rc;
    }
    return MRP_OK;
}

extern "C" int mrp_render_batch(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                                int precision, float* const* dev_out4) {
    if (!c || !vols || !views || !dev_out4 || n_vols < 1 || n_views < 1)
        return mrp_fail(MRP_ERR_INVALID, "bad batch arguments");
    if (precision != MRP_FP32 && precision != MRP_BF16 && precision != MRP_FP8 && precision != MRP_BF16_1T)
        return mrp_fail(MRP_ERR_INVALID, "unknown precision %d", precision);
    FusedState* f = c->fused;
    if (!c->lut_valid || c->lut_g != views[0].g) return mrp_fail(MRP_ERR_STATE, "HG LUT not prepared for g");
    int n_tiles = 0;
    size_t pixels = 0;
    for (int i = 0; i < n_views; ++i) {
        const mrp_view& v = views[i];
        int rc = mrp_check_view(&v);
        if (rc) return rc;
        if (!dev_out4[i]) return mrp_fail(MRP_ERR_INVALID, "view %d: null output", i);
        if (v.volume_index < 0 || v.volume_index >= n_vols || !vols[v.volume_index])
            return mrp_fail(MRP_ERR_INVALID, "view %d: bad volume_index", i);
        if (v.g != views[0].g) return mrp_fail(MRP_ERR_INVALID, "all views of a batch must share g");
        const mrp_volume* vol = vols[v.volume_index];
        if (vol->ctx != c) return mrp_fail(MRP_ERR_INVALID, "volume belongs to another context");
        if (!vol->mips_valid || !vol->tr_valid || vol->tr_alpha != v.alpha ||
            memcmp(vol->tr_light, v.light_dir, sizeof vol->tr_light) != 0)
            return mrp_fail(MRP_ERR_STATE, "view %d: TR mips not prepared (call mrp_prepare_views)", i);
        n_tiles += ((v.width + 7) / 8) * ((v.height + 7) / 8);
        pixels += (size_t)v.width * v.height;
    }
    // staging buffers are reused: wait for the previous batch's uploads
    if (hipEventSynchronize(f->staged) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event sync failed");
    if (n_vols > f->vols_cap) {
        hipHostFree(f->h_vols);
        f->h_vols = nullptr;
        if (hipHostMalloc(&f->h_vols, n_vols * sizeof(DevVolume)) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "pinned allocation failed");
        int rc = grow((void**)&f->d_vols, n_vols * sizeof(DevVolume));
        if (rc) return rc;
        f->vols_cap = n_vols;
    }
    if (n_views > f->slots_cap) {
        hipHostFree(f->h_slots);
        f->h_slots = nullptr;
        if (hipHostMalloc(&f->h_slots, n_views * sizeof(ViewSlot)) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "pinned allocation failed");
        int rc = grow((void**)&f->d_slots, n_views * sizeof(ViewSlot));
        if (rc) return rc;
        f->slots_cap = n_views;
    }
    // each 8x8 tile takes its active pixels rounded up to 16 records: at most 64 per tile
    const size_t need = (size_t)n_tiles * 64;
    if (need > f->cap) {
        int rc = grow((void**)&f->d_recs, need * sizeof(ActiveRec));
        if (rc) return rc;
        f->cap = need;
    }
    for (int i = 0; i < n_vols; ++i) f->h_vols[i] = mrp_dev_volume(vols[i]);
    int t0 = 0;
    for (int i = 0; i < n_views; ++i) {
        ViewSlot& s = f->h_slots[i];
        s.vw = mrp_to_dev(&views[i]);
        s.out = (float4*)dev_out4[i];
        s.vol = views[i].volume_index;
        s.tile0 = t0;
        s.tiles_x = (views[i].width + 7) / 8;
        s.pad = 0;
        t0 += s.tiles_x * ((views[i].height + 7) / 8);
    }
    hipStream_t st = c->stream;
#define HT(x)                                                                               \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return mrp_fail(MRP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
    HT(hipMemcpyAsync(f->d_vols, f->h_vols, n_vols * sizeof(DevVolume), hipMemcpyHostToDevice, st));
    HT(hipMemcpyAsync(f->d_slots, f->h_slots, n_views * sizeof(ViewSlot), hipMemcpyHostToDevice, st));
    HT(hipEventRecord(f->staged, st));
    hipEvent_t* ev = nullptr;
    if (c->timing) {
        int rc = mrp_timing_begin(c, &ev);
        if (rc) return rc;
        HT(hipEventRecord(ev[0], st));
    }
    HT(launch_reset_counters(f->d_counter, f->d_taps, st));
    HT(launch_march_compact(f->d_vols, f->d_slots, n_views, n_tiles, f->d_recs, f->d_counter, f->d_taps, st));
    if (c->timing) HT(hipEventRecord(ev[1], st));
    if (precision == MRP_FP32) {
        HT(launch_predict_active_fp32(f->d_vols, f->d_slots, c->d_stencil, c->d_weights, f->d_recs, f->d_counter,
                                      (int)need, st));
    } else {
        int grid = c->props.multiProcessorCount * f->mfma_waves_per_cu / 4;
        HT(launch_mlp_mfma_active(precision, !c->exact_gather, f->net, f->d_vols, f->d_slots, c->d_stencil, c->d_fast,
                                  c->d_weights, f->d_recs, f->d_counter, grid, st));
    }
    if (c->timing) HT(hipEventRecord(ev[2], st));
    f->have_counts = true;
    c->stats.pixels = (int64_t)pixels;
    c->stats.active_pixels = -1;
    c->stats.march_taps = -1;
#undef HT
    return MRP_OK;
}

int mrp_fused_predict_features(mrp_ctx* c, const float* dev_feats, int n, const float sr[3], float g, int precision,
                               float* dev_rgb) {
    if (precision != MRP_BF16 && precision != MRP_FP8 && precision != MRP_BF16_1T)
        return mrp_fail(MRP_ERR_INVALID, "unknown precision %d", precision);
    hipError_t e = launch_mlp_mfma_features(precision, c->fused->net, c->d_weights, dev_feats, n, sr, g, dev_rgb,
                                            c->stream);
    if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "mfma predictor: %s", hipGetErrorString(e));
    return MRP_OK;
}

extern "C" int mrp_ctx_set_option(mrp_ctx* c, int option, int value) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "ctx is null");
    switch (option) {
        case MRP_OPT_MFMA_WAVES_PER_CU:
            if (value < 1 || value > 32) return mrp_fail(MRP_ERR_INVALID, "MRP_OPT_MFMA_WAVES_PER_CU in [1, 32]");
            c->fused->mfma_waves_per_cu = value;
            return MRP_OK;
        case MRP_OPT_EXACT_GATHER:
            c->exact_gather = value != 0;
            return MRP_OK;
        default:
            return mrp_fail(MRP_ERR_INVALID, "unknown option %d", option);
    }
}

extern "C" int mrp_ctx_get_option(mrp_ctx* c, int option, int* value) {
    if (!c || !value) return mrp_fail(MRP_ERR_INVALID, "null argument");
    switch (option) {
        case MRP_OPT_MFMA_WAVES_PER_CU: *value = c->fused->mfma_waves_per_cu; return MRP_OK;
        case MRP_OPT_EXACT_GATHER: *value = c->exact_gather ? 1 : 0; return MRP_OK;
        default: return mrp_fail(MRP_ERR_INVALID, "unknown option %d", option);
    }
}
