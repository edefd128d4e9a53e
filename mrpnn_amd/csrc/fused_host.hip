// Host side of the batched predictor: view tables, active-list buffers, MFMA weight
// packing, dispatch, statistics.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "capi_internal.h"
#include "fused.h"
#include "mlp_mfma.h"
#include "fp8_scales.h"
#include "kernels.h"

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
    const unsigned* count_src = nullptr;  // counters of the last batch (context or no-alloc buffer)
    const unsigned long long* taps_src = nullptr;
    MfmaNet net;  // packed BF16 / FP8 weights
    int fp8_wexp[ML_COUNT] = {};  // FP8 weight and activation exponents of the layers
    int fp8_aexp[ML_COUNT] = {};
    int mfma_waves_per_cu = 12;  // three 256-thread blocks per CU (three waves per SIMD)
    int tile_lw = 3;             // march tiles of (1 << tile_lw) x (64 >> tile_lw) pixels (8x8)
    bool lds_weights = true;     // one 768-thread block per CU with the weights in LDS
    // graph mode: the batch (staging copies, counter reset, march, network) is captured once
    // per launch shape on a private stream and replayed on the context's stream
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    struct Graph {
        int n_views, n_entries, n_tiles, precision, fast, grid, lds_w, tile_lw;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    // TR sets of the batched path are baked on a second stream: the march of a batch (which
    // does not read TR mips) overlaps the bakes of its suns; the network waits for tr_done
    hipStream_t aux = nullptr;
    hipEvent_t pre = nullptr, tr_done = nullptr;
    bool tr_recorded = false;
};

void mrp::view_const_init(const ViewDev& vw, ViewConst* c) {
    c->L = normalize3(vw.light_dir);
    c->LX = normalize3(c->L);
    c->invLX = inv3(c->LX);
    c->sr = vw.scatter_rate / 1.001f;
    c->sp = mk3(powf(c->sr.x, 4.0f), powf(c->sr.y, 4.0f), powf(c->sr.z, 4.0f));
    c->alpha64 = vw.alpha / 64.0f;
    c->need_y = c->sr.y != c->sr.x;
    c->need_z = c->sr.z != c->sr.x && c->sr.z != c->sr.y;
}

static void clear_graphs(FusedState* f) {
    for (auto& g : f->graphs) hipGraphExecDestroy(g.exec);
    f->graphs.clear();
}

static int grow(void** p, size_t bytes) {
    hipFree(*p);  // callers drop captured graphs first: they hold the old address
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) return mrp_fail(MRP_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
    return MRP_OK;
}

int mrp_fused_init(mrp_ctx* c) {
    FusedState* f = new FusedState();
    c->fused = f;
    if (hipMalloc(&f->d_counter, 16) != hipSuccess || hipMalloc(&f->d_taps, 16) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "counter allocation failed");
    if (hipEventCreateWithFlags(&f->staged, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&f->pre, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&f->tr_done, hipEventDisableTiming) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "event creation failed");
    if (hipStreamCreateWithFlags(&f->aux, hipStreamNonBlocking) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "stream creation failed");
    int rc = mfma_net_create(c->weights_host.data(), &f->net, f->fp8_wexp);
    if (rc) return rc;
    memcpy(f->fp8_aexp, kFp8ActExp, sizeof f->fp8_aexp);
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
    if (f->pre) hipEventDestroy(f->pre);
    if (f->tr_done) hipEventDestroy(f->tr_done);
    if (f->aux) {
        hipStreamSynchronize(f->aux);
        hipStreamDestroy(f->aux);
    }
    clear_graphs(f);
    if (f->cap_stream) hipStreamDestroy(f->cap_stream);
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
        if (hipMemcpyAsync(&cnt, f->count_src, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(&taps, f->taps_src, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return mrp_fail(MRP_ERR_HIP, "reading counters failed");
        c->stats.active_pixels = cnt;
        c->stats.march_taps = (int64_t)taps;
        f->have_counts = false;
    }
    return MRP_OK;
}

// Orders the context's stream after the TR bakes issued so far on the second stream (before
// a volume's pyramid is rewritten or freed).
int mrp_fused_join_aux(mrp_ctx* c) {
    FusedState* f = c->fused;
    if (f && f->tr_recorded && hipStreamWaitEvent(c->stream, f->tr_done, 0) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "stream ordering failed");
    return MRP_OK;
}

// ------------------------------------------------------------------------------------------
// TR cache of the batched path
// ------------------------------------------------------------------------------------------
static bool tr_key_eq(const TrSlot& t, const mrp_volume* v, const float L[3], float alpha) {
    return t.valid && t.vol == v && t.light[0] == L[0] && t.light[1] == L[1] && t.light[2] == L[2] &&
           t.alpha == alpha;
}
int mrp_tr_slot_find(const mrp_ctx* c, const mrp_volume* v, const float L[3], float alpha) {
    for (size_t i = 0; i < c->tr_slots.size(); ++i)
        if (tr_key_eq(c->tr_slots[i], v, L, alpha)) return (int)i;
    return -1;
}
void mrp_tr_slots_drop(mrp_ctx* c, const mrp_volume* v) {
    for (TrSlot& t : c->tr_slots)
        if (t.vol == v) t.valid = false;
}
// A slot for a new key: an unused one, a new one below the cap, else the least recently used
// slot that the current prepare call has not claimed.
static int tr_slot_claim(mrp_ctx* c) {
    int best = -1;
    for (size_t i = 0; i < c->tr_slots.size(); ++i) {
        const TrSlot& t = c->tr_slots[i];
        if (!t.valid) return (int)i;
        if (t.stamp != c->tr_stamp && (best < 0 || t.stamp < c->tr_slots[best].stamp)) best = (int)i;
    }
    if ((int)c->tr_slots.size() < c->tr_slot_cap) {
        TrSlot t;
        if (hipMalloc(&t.d_tr, kTrFloats * sizeof(float)) != hipSuccess ||
            hipMalloc(&t.d_ptr, kPadTrAlloc * sizeof(float)) != hipSuccess ||
            hipEventCreateWithFlags(&t.last_read, hipEventDisableTiming) != hipSuccess) {
            hipFree(t.d_tr);
            hipFree(t.d_ptr);
            return -1;
        }
        // the padded copy's border must read as zero; the bake writes the interior only
        if (hipMemsetAsync(t.d_ptr, 0, kPadTrAlloc * sizeof(float), c->fused->aux) != hipSuccess) return -1;
        c->tr_slots.push_back(t);
        return (int)c->tr_slots.size() - 1;
    }
    return best;
}

// LUT for the batch's g, and a TR set for every distinct (volume, light, alpha) of the views
// (Update_TR, volume.cu:1589-1667, once per key).
extern "C" int mrp_prepare_views(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views) {
    if (!c || !vols || !views || n_vols < 1 || n_views < 1) return mrp_fail(MRP_ERR_INVALID, "bad batch arguments");
    for (int i = 0; i < n_views; ++i) {
        int rc = mrp_check_view(&views[i]);
        if (rc) return rc;
        int vi = views[i].volume_index;
        if (vi < 0 || vi >= n_vols || !vols[vi]) return mrp_fail(MRP_ERR_INVALID, "view %d: bad volume_index", i);
        if (vols[vi]->ctx != c) return mrp_fail(MRP_ERR_INVALID, "view %d: volume belongs to another context", i);
        if (!vols[vi]->mips_valid) return mrp_fail(MRP_ERR_STATE, "view %d: mrp_volume_update must run first", i);
        if (views[i].g != views[0].g) return mrp_fail(MRP_ERR_INVALID, "all views of a batch must share g");
    }
    int rc = mrp_ctx_update_hg_lut(c, views[0].g, c->lut_valid ? c->lut_seed : 1u);
    if (rc) return rc;
    ++c->tr_stamp;
    // a bake on the second stream waits only for what it depends on: the pyramid it reads
    // (mips_ready of the volume) and the last batch that read the set it overwrites
    // (last_read of the slot); earlier batches may still run meanwhile
    FusedState* f = c->fused;
    for (int i = 0; i < n_views; ++i) {
        mrp_volume* v = vols[views[i].volume_index];
        const float* L = views[i].light_dir;
        int s = mrp_tr_slot_find(c, v, L, views[i].alpha);
        if (s < 0) {
            s = tr_slot_claim(c);
            if (s < 0)
                return mrp_fail(MRP_ERR_NOMEM, "TR cache: no free set for view %d (raise MRP_OPT_TR_SLOTS, now %d)", i,
                                c->tr_slot_cap);
            TrSlot& t = c->tr_slots[s];
            t.valid = false;
            if ((t.read_pending && hipStreamWaitEvent(f->aux, t.last_read, 0) != hipSuccess) ||
                (v->mips_ready && hipStreamWaitEvent(f->aux, v->mips_ready, 0) != hipSuccess))
                return mrp_fail(MRP_ERR_HIP, "stream ordering failed");
            rc = mrp_tr_bake_into(v, L, views[i].alpha, t.d_tr, t.d_ptr, f->aux);
            if (rc) return rc;
            t.vol = v;
            memcpy(t.light, L, sizeof t.light);
            t.alpha = views[i].alpha;
            t.valid = true;
        }
        c->tr_slots[s].stamp = c->tr_stamp;
    }
    if (hipEventRecord(f->tr_done, f->aux) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event record failed");
    f->tr_recorded = true;
    return MRP_OK;
}

extern "C" int mrp_render_batch(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                                int precision, float* const* dev_out4) {
    return mrp_render_batch_checker(c, vols, n_vols, views, n_views, precision, dev_out4, nullptr);
}

// Tiles of a view with march tiles (1 << lw) pixels wide and (64 >> lw) high.
static inline int tiles_x_of(int width, int lw) { return (width + (1 << lw) - 1) >> lw; }
static inline int tiles_y_of(int height, int lw) { return (height + (64 >> lw) - 1) / (64 >> lw); }

// Checks a batch's arguments and finds the TR set of every view (view_slot[i]); counts its
// 64-pixel march tiles and pixels.
static int batch_validate(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                          int precision, float* const* dev_out4, int* view_slot, int* n_tiles, size_t* pixels) {
    const int lw = c->fused->tile_lw;
    if (!c || !vols || !views || !dev_out4 || n_vols < 1 || n_views < 1)
        return mrp_fail(MRP_ERR_INVALID, "bad batch arguments");
    if (precision != MRP_FP32 && precision != MRP_BF16 && precision != MRP_FP8 && precision != MRP_BF16_1T)
        return mrp_fail(MRP_ERR_INVALID, "unknown precision %d", precision);
    if (!c->lut_valid || c->lut_g != views[0].g) return mrp_fail(MRP_ERR_STATE, "HG LUT not prepared for g");
    *n_tiles = 0;
    *pixels = 0;
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
        const int slot = vol->mips_valid ? mrp_tr_slot_find(c, vol, v.light_dir, v.alpha) : -1;
        if (slot < 0) return mrp_fail(MRP_ERR_STATE, "view %d: TR mips not prepared (call mrp_prepare_views)", i);
        view_slot[i] = slot;
        *n_tiles += tiles_x_of(v.width, lw) * tiles_y_of(v.height, lw);
        *pixels += (size_t)v.width * v.height;
    }
    return MRP_OK;
}

// Host images of a batch's tables: one DevVolume per distinct TR set of the views (entry_slot[e]
// is the set of entry e), one ViewSlot per view.  Returns the number of entries.
static int batch_tables(const mrp_ctx* c, mrp_volume* const* vols, const mrp_view* views, int n_views,
                        const int* view_slot, const int* checker, float* const* dev_out4, DevVolume* hv, ViewSlot* hs,
                        int* entry_slot) {
    int n_entries = 0;
    int t0 = 0;
    const int lw = c->fused->tile_lw;
    for (int i = 0; i < n_views; ++i) {
        const int sl = view_slot[i];
        int e = 0;
        while (e < n_entries && entry_slot[e] != sl) ++e;
        if (e == n_entries) {
            const TrSlot& t = c->tr_slots[sl];
            DevVolume d = mrp_dev_volume(vols[views[i].volume_index]);
            size_t o = 0;
            for (int m = 0; m < 8; ++m) {
                d.tr[m] = t.d_tr + o;
                o += (size_t)(128 >> m) * (128 >> m) * (128 >> m);
            }
            d.ptr = t.d_ptr;
            entry_slot[n_entries] = sl;
            hv[n_entries++] = d;
        }
        ViewSlot& s = hs[i];
        s.vw = mrp_to_dev(&views[i]);
        view_const_init(s.vw, &s.vc);
        s.out = (float4*)dev_out4[i];
        s.vol = e;
        s.tile0 = t0;
        s.tiles_x = tiles_x_of(views[i].width, lw);
        s.checker = checker ? checker[i] : 0;
        s.tile_lw = lw;
        t0 += s.tiles_x * tiles_y_of(views[i].height, lw);
    }
    return n_entries;
}

// The TR sets a batch reads may be re-baked (by a later mrp_prepare_views on the second stream)
// only after it: one event per set, created when the set was claimed.
static int batch_mark_reads(mrp_ctx* c, const int* entry_slot, int n_entries, hipStream_t st) {
    for (int e = 0; e < n_entries; ++e) {
        TrSlot& t = c->tr_slots[entry_slot[e]];
        if (!t.last_read || hipEventRecord(t.last_read, st) != hipSuccess)
            return mrp_fail(MRP_ERR_HIP, "recording the TR set's read event failed");
        t.read_pending = true;
    }
    return MRP_OK;
}

// Network launch shape: the LDS-weight kernel runs one 768-thread workgroup per CU, the
// L1/L2-weight kernel mfma_waves_per_cu / 4 256-thread workgroups per CU.
static hipError_t launch_network(const mrp_ctx* c, int precision, const DevVolume* d_vols, const ViewSlot* d_slots,
                                 const ActiveRec* d_recs, const unsigned* d_counter, size_t capacity, hipStream_t q) {
    const FusedState* f = c->fused;
    if (precision == MRP_FP32)
        return launch_predict_active_fp32(d_vols, d_slots, c->d_stencil, c->d_weights, d_recs, d_counter,
                                          (int)capacity, q);
    const bool lds_w = f->lds_weights;
    const int grid = lds_w ? c->props.multiProcessorCount : c->props.multiProcessorCount * f->mfma_waves_per_cu / 4;
    return launch_mlp_mfma_active(precision, !c->exact_gather, lds_w, f->net, d_vols, d_slots, c->d_stencil,
                                  c->d_fast, c->d_lut_rows, c->d_weights, d_recs, d_counter, grid, q);
}

int mrp_render_batch_checker(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                             int precision, float* const* dev_out4, const int* checker) {
    if (!c || n_views < 1) return mrp_fail(MRP_ERR_INVALID, "bad batch arguments");
    FusedState* f = c->fused;
    std::vector<int> view_slot(n_views, -1);
    int n_tiles = 0;
    size_t pixels = 0;
    int vrc = batch_validate(c, vols, n_vols, views, n_views, precision, dev_out4, view_slot.data(), &n_tiles, &pixels);
    if (vrc) return vrc;
    // staging buffers are reused: wait for the previous batch's uploads
    if (hipEventSynchronize(f->staged) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event sync failed");
    // one DevVolume entry per distinct (volume, TR set) of the batch
    if (n_views > f->vols_cap) {
        clear_graphs(f);
        hipHostFree(f->h_vols);
        f->h_vols = nullptr;
        if (hipHostMalloc(&f->h_vols, n_views * sizeof(DevVolume)) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "pinned allocation failed");
        clear_graphs(f);
        int rc = grow((void**)&f->d_vols, n_views * sizeof(DevVolume));
        if (rc) return rc;
        f->vols_cap = n_views;
    }
    if (n_views > f->slots_cap) {
        clear_graphs(f);
        hipHostFree(f->h_slots);
        f->h_slots = nullptr;
        if (hipHostMalloc(&f->h_slots, n_views * sizeof(ViewSlot)) != hipSuccess)
            return mrp_fail(MRP_ERR_NOMEM, "pinned allocation failed");
        clear_graphs(f);
        int rc = grow((void**)&f->d_slots, n_views * sizeof(ViewSlot));
        if (rc) return rc;
        f->slots_cap = n_views;
    }
    // each 64-pixel tile takes its active pixels rounded up to 16 records: at most 64 per tile
    const size_t need = (size_t)n_tiles * 64;
    if (need > f->cap) {
        clear_graphs(f);
        int rc = grow((void**)&f->d_recs, need * sizeof(ActiveRec));
        if (rc) return rc;
        f->cap = need;
    }
    std::vector<int> entry_slot(n_views);
    const int n_entries =
        batch_tables(c, vols, views, n_views, view_slot.data(), checker, dev_out4, f->h_vols, f->h_slots, entry_slot.data());
    hipStream_t st = c->stream;
#define HT(x)                                                                               \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return mrp_fail(MRP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
    // LDS-weight kernel: one 768-thread workgroup per CU; else 256-thread workgroups
    const bool lds_w = f->lds_weights && precision != MRP_FP32;
    const int grid = lds_w ? c->props.multiProcessorCount : c->props.multiProcessorCount * f->mfma_waves_per_cu / 4;
    const bool fast = !c->exact_gather;
    // the work of one batch on stream q; ev (optional) brackets march and network
    auto enqueue = [&](hipStream_t q, hipEvent_t* ev, hipEvent_t staged, bool wait_tr) -> hipError_t {
        hipError_t e;
        if ((e = hipMemcpyAsync(f->d_vols, f->h_vols, n_entries * sizeof(DevVolume), hipMemcpyHostToDevice, q)))
            return e;
        if ((e = hipMemcpyAsync(f->d_slots, f->h_slots, n_views * sizeof(ViewSlot), hipMemcpyHostToDevice, q)))
            return e;
        if (staged && (e = hipEventRecord(staged, q))) return e;
        if (ev && (e = hipEventRecord(ev[0], q))) return e;
        if ((e = launch_reset_counters(f->d_counter, f->d_taps, q))) return e;
        if ((e = launch_march_compact(f->d_vols, f->d_slots, n_views, n_tiles, f->d_recs, f->d_counter, f->d_taps, q)))
            return e;
        if (ev && (e = hipEventRecord(ev[1], q))) return e;
        if (wait_tr && (e = hipStreamWaitEvent(q, f->tr_done, 0))) return e;
        if (ev && (e = hipEventRecord(ev[3], q))) return e;
        if ((e = launch_network(c, precision, f->d_vols, f->d_slots, f->d_recs, f->d_counter, need, q))) return e;
        if (ev && (e = hipEventRecord(ev[2], q))) return e;
        return hipSuccess;
    };
    hipEvent_t* ev = nullptr;
    if (c->timing) {
        int rc = mrp_timing_begin(c, &ev);
        if (rc) return rc;
    }
    if (!f->use_graph) {
        HT(enqueue(st, ev, f->staged, f->tr_recorded));
    } else {
        FusedState::Graph* g = nullptr;
        for (auto& x : f->graphs)
            if (x.n_views == n_views && x.n_entries == n_entries && x.n_tiles == n_tiles && x.precision == precision &&
                x.fast == (int)fast && x.grid == grid && x.lds_w == (int)lds_w && x.tile_lw == f->tile_lw)
                g = &x;
        if (!g) {
            if (!f->cap_stream) HT(hipStreamCreateWithFlags(&f->cap_stream, hipStreamNonBlocking));
            if (f->graphs.size() >= 16) clear_graphs(f);
            hipGraph_t graph = nullptr;
            HT(hipStreamBeginCapture(f->cap_stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = enqueue(f->cap_stream, nullptr, nullptr, false);
            hipError_t ee = hipStreamEndCapture(f->cap_stream, &graph);
            HT(ce);
            HT(ee);
            hipGraphExec_t exec = nullptr;
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            HT(ie);
            f->graphs.push_back({n_views, n_entries, n_tiles, precision, (int)fast, grid, (int)lds_w, f->tile_lw, exec});
            g = &f->graphs.back();
        }
        // graph mode times the whole replay as the network time (march time reads 0)
        if (f->tr_recorded) HT(hipStreamWaitEvent(st, f->tr_done, 0));
        if (ev) {
            HT(hipEventRecord(ev[0], st));
            HT(hipEventRecord(ev[1], st));
            HT(hipEventRecord(ev[3], st));
        }
        HT(hipGraphLaunch(g->exec, st));
        if (ev) HT(hipEventRecord(ev[2], st));
        HT(hipEventRecord(f->staged, st));
    }
    int mrc = batch_mark_reads(c, entry_slot.data(), n_entries, st);
    if (mrc) return mrc;
    f->count_src = f->d_counter;
    f->taps_src = f->d_taps;
    f->have_counts = true;
    c->stats.pixels = (int64_t)pixels;
    c->stats.active_pixels = -1;
    c->stats.march_taps = -1;
#undef HT
    return MRP_OK;
}

// Tooling: the fast gather's features of one view.  The batch (march, compaction, network) runs
// once into a temporary image; the features kernel then reads the same active records.
extern "C" int mrp_gather_features_fast(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, float* dev_feats) {
    if (!c || !v || !vw || !dev_feats) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    mrp_view view = *vw;
    view.volume_index = 0;
    mrp_volume* vols[1] = {v};
    if ((rc = mrp_prepare_views(c, vols, 1, &view, 1))) return rc;
    float* tmp = nullptr;
    if (hipMalloc(&tmp, (size_t)view.width * view.height * 4 * sizeof(float)) != hipSuccess)
        return mrp_fail(MRP_ERR_NOMEM, "temporary image allocation failed");
    const bool graph = c->fused->use_graph;
    c->fused->use_graph = false;  // the features kernel reads the context's own batch buffers
    rc = mrp_render_batch(c, vols, 1, &view, 1, MRP_BF16, &tmp);
    c->fused->use_graph = graph;
    if (!rc) {
        const FusedState* f = c->fused;
        const hipError_t e = launch_fast_features(f->d_vols, f->d_slots, c->d_fast, c->d_lut_rows, f->d_recs,
                                                  f->d_counter, c->props.multiProcessorCount * 4, dev_feats, c->stream);
        if (e != hipSuccess) rc = mrp_fail(MRP_ERR_HIP, "fast features: %s", hipGetErrorString(e));
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = mrp_fail(MRP_ERR_HIP, "stream sync failed");
    (void)hipFree(tmp);
    return rc;
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

// ------------------------------------------------------------------------------------------
// FP8 activation scales
// ------------------------------------------------------------------------------------------
static_assert(MRP_FP8_LAYERS == ML_COUNT, "FP8 layer count of the C ABI");

extern "C" int mrp_fp8_calibrate(mrp_ctx* c, const float* dev_feats, int n, const float sr[3], float g, int reset,
                                 float* maxima) {
    if (!c || !sr || (n > 0 && !dev_feats)) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (n < 0) return mrp_fail(MRP_ERR_INVALID, "n must be >= 0");
    FusedState* f = c->fused;
    const MfmaNet& net = f->net;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "stream sync failed");
    if (reset && hipMemsetAsync(mfma_cal_words(net), 0, ML_COUNT * 4, c->stream) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "clearing the calibration maxima failed");
    if (n > 0) {
        hipError_t e = launch_mlp_mfma_calibrate(net, c->d_weights, dev_feats, n, sr, g, c->stream);
        if (e != hipSuccess) return mrp_fail(MRP_ERR_HIP, "calibration kernel: %s", hipGetErrorString(e));
    }
    unsigned bits[ML_COUNT];
    if (hipMemcpyAsync(bits, mfma_cal_words(net), sizeof bits, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "reading the calibration maxima failed");
    int exps[ML_COUNT];
    for (int id = 0; id < ML_COUNT; ++id) {
        float m;
        memcpy(&m, &bits[id], 4);
        if (!(m >= 0.0f) || !isfinite(m)) return mrp_fail(MRP_ERR_STATE, "layer %d: non-finite activation maximum", id);
        exps[id] = fp8_exp_for_max(m);
        if (maxima) maxima[id] = m;
    }
    int rc = mfma_net_set_fp8_exps(net, f->fp8_wexp, exps);
    if (rc) return rc;
    memcpy(f->fp8_aexp, exps, sizeof exps);
    return MRP_OK;
}

extern "C" int mrp_fp8_get_exponents(mrp_ctx* c, int* exps, int n) {
    if (!c || !exps) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (n != ML_COUNT) return mrp_fail(MRP_ERR_INVALID, "n must be MRP_FP8_LAYERS (%d)", ML_COUNT);
    memcpy(exps, c->fused->fp8_aexp, sizeof(int) * ML_COUNT);
    return MRP_OK;
}

extern "C" int mrp_fp8_set_exponents(mrp_ctx* c, const int* exps, int n) {
    if (!c || !exps) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (n != ML_COUNT) return mrp_fail(MRP_ERR_INVALID, "n must be MRP_FP8_LAYERS (%d)", ML_COUNT);
    for (int id = 0; id < ML_COUNT; ++id)
        if (exps[id] < -100 || exps[id] > 100) return mrp_fail(MRP_ERR_INVALID, "exponent of layer %d out of range", id);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "stream sync failed");
    FusedState* f = c->fused;
    int rc = mfma_net_set_fp8_exps(f->net, f->fp8_wexp, exps);
    if (rc) return rc;
    memcpy(f->fp8_aexp, exps, sizeof(int) * ML_COUNT);
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
        case MRP_OPT_GRAPH:
            c->fused->use_graph = value != 0;
            return MRP_OK;
        case MRP_OPT_LDS_WEIGHTS:
            c->fused->lds_weights = value != 0;
            return MRP_OK;
        case MRP_OPT_TILE_W: {
            int lw = 2;
            while (lw < 6 && (1 << lw) != value) ++lw;
            if ((1 << lw) != value) return mrp_fail(MRP_ERR_INVALID, "MRP_OPT_TILE_W is 4, 8, 16, 32 or 64");
            c->fused->tile_lw = lw;
            return MRP_OK;
        }
        case MRP_OPT_RPNN_CHUNK:
            if (value < 4096 || value > 1048576 || value % 128)
                return mrp_fail(MRP_ERR_INVALID, "MRP_OPT_RPNN_CHUNK in [4096, 1048576], a multiple of 128");
            c->rpnn_chunk = value;
            return MRP_OK;
        case MRP_OPT_TR_SLOTS:
            if (value < 1 || value > 4096) return mrp_fail(MRP_ERR_INVALID, "MRP_OPT_TR_SLOTS in [1, 4096]");
            if (value < (int)c->tr_slots.size()) {  // shrink: free the tail sets
                if (hipStreamSynchronize(c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "sync failed");
                if (hipStreamSynchronize(c->fused->aux) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "sync failed");
                for (size_t i = value; i < c->tr_slots.size(); ++i) {
                    hipFree(c->tr_slots[i].d_tr);
                    hipFree(c->tr_slots[i].d_ptr);
                    if (c->tr_slots[i].last_read) hipEventDestroy(c->tr_slots[i].last_read);
                }
                c->tr_slots.resize(value);
            }
            c->tr_slot_cap = value;
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
        case MRP_OPT_TR_SLOTS: *value = c->tr_slot_cap; return MRP_OK;
        case MRP_OPT_GRAPH: *value = c->fused->use_graph ? 1 : 0; return MRP_OK;
        case MRP_OPT_LDS_WEIGHTS: *value = c->fused->lds_weights ? 1 : 0; return MRP_OK;
        case MRP_OPT_TILE_W: *value = 1 << c->fused->tile_lw; return MRP_OK;
        case MRP_OPT_RPNN_CHUNK: *value = c->rpnn_chunk; return MRP_OK;
        default: return mrp_fail(MRP_ERR_INVALID, "unknown option %d", option);
    }
}

// Launch-shape autotuner: times the batch (already prepared) with each candidate number of
// persistent waves per CU of the network kernel and keeps the fastest.
extern "C" int mrp_autotune_batch(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views, int n_views,
                                  int precision, float* const* dev_out4, int* best_waves, float* best_ms) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "ctx is null");
    FusedState* f = c->fused;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return mrp_fail(MRP_ERR_HIP, "event creation failed");
    // milliseconds per batch of the current launch shape: one warm-up, then three timed batches
    auto time_batch = [&](float* ms) -> int {
        int rc = mrp_render_batch(c, vols, n_vols, views, n_views, precision, dev_out4);
        if (rc) return rc;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event record failed");
        for (int r = 0; r < 3 && !rc; ++r) rc = mrp_render_batch(c, vols, n_vols, views, n_views, precision, dev_out4);
        if (rc) return rc;
        float t = 0.0f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            return mrp_fail(MRP_ERR_HIP, "event timing failed");
        *ms = t / 3.0f;
        return MRP_OK;
    };
    const int keep = f->mfma_waves_per_cu, keep_lw = f->tile_lw;
    const bool keep_lds = f->lds_weights;
    int rc = MRP_OK;
    float best_t = 1e30f;
    // 1. the march tile shape (widths 8, 16, 4, 32, 64), with the network's current launch
    int best_lw = keep_lw;
    static const int lws[] = {3, 4, 2, 5, 6};
    for (int lw : lws) {
        f->tile_lw = lw;
        float ms = 0.0f;
        if ((rc = time_batch(&ms))) break;
        if (ms < best_t) {
            best_t = ms;
            best_lw = lw;
        }
    }
    f->tile_lw = rc ? keep_lw : best_lw;
    // 2. the network's launch: LDS-resident weights (one 768-thread workgroup per CU), then
    // L1/L2 weights at each waves-per-CU count
    static const int cand[] = {0, 4, 8, 12, 16, 24, 32};
    int best = keep;
    bool best_lds = keep_lds;
    best_t = 1e30f;
    for (int w : cand) {
        if (rc) break;
        f->lds_weights = w == 0;
        f->mfma_waves_per_cu = w == 0 ? keep : w;
        float ms = 0.0f;
        if ((rc = time_batch(&ms))) break;
        if (ms < best_t) {
            best_t = ms;
            best = w == 0 ? keep : w;
            best_lds = w == 0;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    f->mfma_waves_per_cu = rc ? keep : best;
    f->lds_weights = rc ? keep_lds : best_lds;
    if (rc) {
        f->tile_lw = keep_lw;
        return rc;
    }
    if (best_waves) *best_waves = best_lds ? 12 : best;
    if (best_ms) *best_ms = best_t;
    return MRP_OK;
}

// ------------------------------------------------------------------------------------------
// mrp_render_batch_noalloc: mrp_render_batch with its per-call device buffers in one static
// device array, for callers that cannot allocate.  Layout of the array (bytes):
//   [0, 4)            record counter of the march (unsigned)
//   [128, 136)        march taps (unsigned long long)
//   [256, ...)        DevVolume table, MRP_NOALLOC_MAX_VIEWS entries
//   [kNaSlots, ...)   ViewSlot table, MRP_NOALLOC_MAX_VIEWS entries
//   [65536, end)      active-pixel records, MRP_NOALLOC_MAX_RECORDS x 64 bytes
// ------------------------------------------------------------------------------------------
namespace {
constexpr size_t kNaCounter = 0;
constexpr size_t kNaTaps = 128;
constexpr size_t kNaVols = 256;
constexpr size_t kNaSlots = kNaVols + ((MRP_NOALLOC_MAX_VIEWS * sizeof(DevVolume) + 255) & ~(size_t)255);
constexpr size_t kNaRecs = 65536;
static_assert(kNaSlots + MRP_NOALLOC_MAX_VIEWS * sizeof(ViewSlot) <= kNaRecs, "tables overrun the records");
static_assert(kNaRecs + (size_t)MRP_NOALLOC_MAX_RECORDS * sizeof(ActiveRec) == MRP_NOALLOC_SCRATCH_BYTES,
              "MRP_NOALLOC_SCRATCH_BYTES does not match the layout");
constexpr int kNaMaxDevices = 64;
std::mutex na_mu;
unsigned char* na_ptr[kNaMaxDevices];  // device address of `scratch` on each device, once resolved
hipError_t na_err[kNaMaxDevices];
bool na_resolved[kNaMaxDevices];

// Tables travel as kernel arguments (no host staging buffer, no copy from pageable memory):
// up to four volume entries and four view slots per launch.
constexpr int kNaChunk = 4;
struct TableChunk {
    int n_vols, vol0, n_slots, slot0;
    DevVolume v[kNaChunk];
    ViewSlot s[kNaChunk];
};
static_assert(sizeof(TableChunk) <= 2048, "kernel argument block too large");
}  // namespace

// The scratch array.  In host code its name is a host-side stand-in and never the device
// memory: host code uses the address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[MRP_NOALLOC_SCRATCH_BYTES];

__global__ void k_stage_tables(TableChunk t, DevVolume* vols, ViewSlot* slots) {
    const int i = threadIdx.x;
    if (i < t.n_vols) vols[t.vol0 + i] = t.v[i];
    if (i < t.n_slots) slots[t.slot0 + i] = t.s[i];
}

// Once per device, with that device current (called by mrp_ctx_create): resolves the device
// address of the scratch array.  A failure is kept and reported by mrp_render_batch_noalloc.
void mrp_noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return;
    std::lock_guard<std::mutex> lk(na_mu);
    if (na_resolved[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) (void)hipGetLastError();  // do not leave the error for the next launch check
    na_err[device] = e;
    na_ptr[device] = e == hipSuccess ? (unsigned char*)p : nullptr;
    na_resolved[device] = true;
}

extern "C" int mrp_render_batch_noalloc(mrp_ctx* c, mrp_volume* const* vols, int n_vols, const mrp_view* views,
                                        int n_views, int precision, float* const* dev_out4) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "ctx is null");
    if (n_views < 1 || n_views > MRP_NOALLOC_MAX_VIEWS)
        return mrp_fail(MRP_ERR_INVALID, "n_views %d outside [1, %d]", n_views, MRP_NOALLOC_MAX_VIEWS);
    int view_slot[MRP_NOALLOC_MAX_VIEWS];
    int n_tiles = 0;
    size_t pixels = 0;
    int rc = batch_validate(c, vols, n_vols, views, n_views, precision, dev_out4, view_slot, &n_tiles, &pixels);
    if (rc) return rc;
    const size_t need = (size_t)n_tiles * 64;
    if (need > (size_t)MRP_NOALLOC_MAX_RECORDS)
        return mrp_fail(MRP_ERR_INVALID, "the batch needs %zu active-pixel records (64 per march tile), the limit is %d",
                        need, MRP_NOALLOC_MAX_RECORDS);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != c->device)
        return mrp_fail(MRP_ERR_STATE, "the current device (%d) is not the context's (%d)", cur, c->device);
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(na_mu);
        if (c->device >= 0 && c->device < kNaMaxDevices && na_resolved[c->device]) {
            base = na_ptr[c->device];
            if (!base)
                return mrp_fail(MRP_ERR_HIP, "scratch array of device %d unavailable: %s", c->device,
                                hipGetErrorString(na_err[c->device]));
        }
    }
    if (!base) return mrp_fail(MRP_ERR_STATE, "scratch array of device %d not resolved", c->device);
    unsigned* counter = (unsigned*)(base + kNaCounter);
    unsigned long long* taps = (unsigned long long*)(base + kNaTaps);
    DevVolume* d_vols = (DevVolume*)(base + kNaVols);
    ViewSlot* d_slots = (ViewSlot*)(base + kNaSlots);
    ActiveRec* d_recs = (ActiveRec*)(base + kNaRecs);

    DevVolume hv[MRP_NOALLOC_MAX_VIEWS];
    ViewSlot hs[MRP_NOALLOC_MAX_VIEWS];
    int entry_slot[MRP_NOALLOC_MAX_VIEWS];
    const int n_entries = batch_tables(c, vols, views, n_views, view_slot, nullptr, dev_out4, hv, hs, entry_slot);
    FusedState* f = c->fused;
    hipStream_t st = c->stream;
#define HT(x)                                                                                    \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return mrp_fail(MRP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
    for (int k = 0; k < n_views || k < n_entries; k += kNaChunk) {
        TableChunk t;
        t.vol0 = t.slot0 = k;
        t.n_vols = n_entries - k < 0 ? 0 : (n_entries - k < kNaChunk ? n_entries - k : kNaChunk);
        t.n_slots = n_views - k < kNaChunk ? n_views - k : kNaChunk;
        for (int i = 0; i < t.n_vols; ++i) t.v[i] = hv[k + i];
        for (int i = 0; i < t.n_slots; ++i) t.s[i] = hs[k + i];
        hipLaunchKernelGGL(k_stage_tables, dim3(1), dim3(64), 0, st, t, d_vols, d_slots);
        HT(hipGetLastError());
    }
    HT(launch_reset_counters(counter, taps, st));
    HT(launch_march_compact(d_vols, d_slots, n_views, n_tiles, d_recs, counter, taps, st));
    if (f->tr_recorded) HT(hipStreamWaitEvent(st, f->tr_done, 0));
    HT(launch_network(c, precision, d_vols, d_slots, d_recs, counter, need, st));
#undef HT
    rc = batch_mark_reads(c, entry_slot, n_entries, st);
    if (rc) return rc;
    f->count_src = counter;
    f->taps_src = taps;
    f->have_counts = true;
    c->stats.pixels = (int64_t)pixels;
    c->stats.active_pixels = -1;
    c->stats.march_taps = -1;
    return MRP_OK;
}
