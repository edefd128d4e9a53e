// Serving loop of the batched predictor: device-resident volumes admitted by free HBM,
// frame requests batched as they arrive (continuous batching: while one batch runs on the
// GPU the next one is formed from whatever has arrived and launched as soon as an in-flight
// slot frees up), completions reported with submit/launch/done times.
//
// No counterpart in the reference, which renders one view per blocking Render call
// (core/volume.cu:487-548); the plain per-frame path of this library stays the drop-in and
// is available here as the baseline mode (one plain launch per request, FIFO).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/mrp_c_api.h"
#include "capi_internal.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SrvVolume {
    int N = 0;
    float max_density = 1.0f;
    std::vector<float> host;  // kept for (re)admission
    int fbm_octaves = 0;      // > 0: regenerate from fBm instead of a host copy
    float fbm_freq = 4.0f;
    uint32_t fbm_seed = 0;
    mrp_volume* vol = nullptr;  // resident when non-null
    int64_t bytes = 0;
    int pending = 0;            // requests queued or in flight
    uint64_t last_use = 0;
};

struct SrvReq {
    mrp_request r;
    double t_submit = 0.0, t_launch = 0.0;
};

struct InFlight {
    hipEvent_t done = nullptr;
    std::vector<SrvReq> reqs;
};

}  // namespace

struct mrp_server {
    mrp_ctx* ctx = nullptr;
    mrp_server_config cfg{};
    std::vector<SrvVolume> vols;
    std::deque<SrvReq> pending;
    std::deque<InFlight> inflight;
    std::vector<hipEvent_t> free_events;
    std::vector<mrp_completion> done;
    std::mutex mu;  // submit may run on another thread than step
    int64_t resident_bytes = 0;
    uint64_t tick = 0;
    mrp_server_stats stats{};
};

static int srv_free_bytes(mrp_server* s, int64_t* out) {
    if (s->cfg.hbm_limit_bytes > 0) {
        *out = s->cfg.hbm_limit_bytes - s->resident_bytes;
        return MRP_OK;
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "hipMemGetInfo failed");
    *out = (int64_t)fr;
    return MRP_OK;
}

static int srv_evict(mrp_server* s, SrvVolume& v) {
    mrp_volume_destroy(v.vol);
    v.vol = nullptr;
    s->resident_bytes -= v.bytes;
    s->stats.evictions += 1;
    return MRP_OK;
}

// Makes volume i resident if its pyramid fits in free HBM, evicting least recently used idle
// volumes when that makes room.  Returns 1 when resident, 0 when it has to wait.
static int srv_admit(mrp_server* s, int i) {
    SrvVolume& v = s->vols[i];
    if (v.vol) return 1;
    const int64_t need = v.bytes + s->cfg.hbm_reserve_bytes;
    if (s->cfg.hbm_limit_bytes > 0 && need > s->cfg.hbm_limit_bytes)
        return -mrp_fail(MRP_ERR_NOMEM, "volume %d (%lld bytes) exceeds the HBM limit", i, (long long)v.bytes);
    for (;;) {
        int64_t fr = 0;
        int rc = srv_free_bytes(s, &fr);
        if (rc) return -rc;
        if (fr >= need) break;
        int lru = -1;
        for (size_t k = 0; k < s->vols.size(); ++k) {
            const SrvVolume& o = s->vols[k];
            if (o.vol && o.pending == 0 && (lru < 0 || o.last_use < s->vols[lru].last_use)) lru = (int)k;
        }
        if (lru < 0) return 0;  // everything resident is busy: stay queued
        srv_evict(s, s->vols[lru]);
    }
    mrp_volume* vol = nullptr;
    int rc = mrp_volume_create(s->ctx, v.N, &vol);
    if (rc) return -rc;
    if (v.fbm_octaves > 0) rc = mrp_volume_fill_fbm(vol, v.fbm_octaves, v.fbm_freq, v.fbm_seed);
    else rc = mrp_volume_set_data(vol, v.host.data(), v.max_density);
    if (!rc) rc = mrp_volume_update(vol);
    if (rc) {
        mrp_volume_destroy(vol);
        return -rc;
    }
    v.vol = vol;
    s->resident_bytes += v.bytes;
    s->stats.admissions += 1;
    return 1;
}

static void srv_reap(mrp_server* s, bool wait) {
    while (!s->inflight.empty()) {
        InFlight& f = s->inflight.front();
        if (wait) hipEventSynchronize(f.done);
        else if (hipEventQuery(f.done) != hipSuccess) break;
        const double t = now_s();
        for (const SrvReq& q : f.reqs) {
            mrp_completion c;
            c.user_tag = q.r.user_tag;
            c.t_submit = q.t_submit;
            c.t_launch = q.t_launch;
            c.t_done = t;
            c.status = MRP_OK;
            s->done.push_back(c);
            s->vols[q.r.volume_id].pending -= 1;
            s->stats.completed += 1;
            s->stats.pixels += (int64_t)q.r.view.width * q.r.view.height;
        }
        s->free_events.push_back(f.done);
        s->inflight.pop_front();
    }
}

extern "C" int mrp_server_create(mrp_ctx* ctx, const mrp_server_config* cfg, mrp_server** out) {
    if (!ctx || !cfg || !out) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (cfg->max_batch_views < 1 || cfg->max_batch_pixels < 1 || cfg->max_inflight < 1)
        return mrp_fail(MRP_ERR_INVALID, "max_batch_views, max_batch_pixels and max_inflight must be >= 1");
    if (cfg->precision != MRP_FP32 && cfg->precision != MRP_BF16 && cfg->precision != MRP_FP8 &&
        cfg->precision != MRP_BF16_1T)
        return mrp_fail(MRP_ERR_INVALID, "unknown precision %d", cfg->precision);
    mrp_server* s = new mrp_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    *out = s;
    return MRP_OK;
}

extern "C" int mrp_server_destroy(mrp_server* s) {
    if (!s) return MRP_OK;
    srv_reap(s, true);
    for (SrvVolume& v : s->vols)
        if (v.vol) mrp_volume_destroy(v.vol);
    for (hipEvent_t e : s->free_events) hipEventDestroy(e);
    delete s;
    return MRP_OK;
}

static int srv_add(mrp_server* s, SrvVolume&& v, int* id) {
    v.bytes = mrp_volume_bytes(v.N);
    std::lock_guard<std::mutex> lk(s->mu);
    s->vols.push_back(std::move(v));
    if (id) *id = (int)s->vols.size() - 1;
    return MRP_OK;
}

extern "C" int mrp_server_add_volume(mrp_server* s, int N, const float* host_grid, float max_density, int* volume_id) {
    if (!s || !host_grid) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (N < 2 || N > 2048) return mrp_fail(MRP_ERR_INVALID, "resolution must be in [2, 2048]");
    if (!(max_density > 0.0f) || !isfinite(max_density)) return mrp_fail(MRP_ERR_INVALID, "max_density must be positive");
    SrvVolume v;
    v.N = N;
    v.max_density = max_density;
    v.host.assign(host_grid, host_grid + (size_t)N * N * N);
    return srv_add(s, std::move(v), volume_id);
}

extern "C" int mrp_server_add_fbm_volume(mrp_server* s, int N, int octaves, float base_freq, uint32_t seed,
                                         int* volume_id) {
    if (!s) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (N < 2 || N > 2048) return mrp_fail(MRP_ERR_INVALID, "resolution must be in [2, 2048]");
    if (octaves < 1 || octaves > 16) return mrp_fail(MRP_ERR_INVALID, "octaves must be in [1, 16]");
    SrvVolume v;
    v.N = N;
    v.fbm_octaves = octaves;
    v.fbm_freq = base_freq;
    v.fbm_seed = seed;
    return srv_add(s, std::move(v), volume_id);
}

extern "C" int mrp_server_submit(mrp_server* s, const mrp_request* r) {
    if (!s || !r) return mrp_fail(MRP_ERR_INVALID, "null argument");
    int rc = mrp_check_view(&r->view);
    if (rc) return rc;
    if (!r->dev_out4) return mrp_fail(MRP_ERR_INVALID, "null output buffer");
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->volume_id < 0 || r->volume_id >= (int)s->vols.size())
        return mrp_fail(MRP_ERR_INVALID, "unknown volume id %d", r->volume_id);
    if (r->view.g != s->cfg.g) return mrp_fail(MRP_ERR_INVALID, "request g differs from the server's g");
    SrvReq q;
    q.r = *r;
    q.t_submit = now_s();
    s->pending.push_back(q);
    s->vols[r->volume_id].pending += 1;
    s->stats.submitted += 1;
    return MRP_OK;
}

// Batched mode: one prepare + render_batch over up to max_batch_views requests of resident
// volumes (arrival order, skipping requests whose volume waits for memory).
static int srv_launch_batch(mrp_server* s) {
    std::vector<SrvReq> take;
    std::vector<mrp_volume*> vols;
    std::vector<int> vol_slot(s->vols.size(), -1);
    std::vector<mrp_view> views;
    std::vector<float*> outs;
    int64_t px = 0;
    int keys = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            if ((int)take.size() >= s->cfg.max_batch_views || px >= s->cfg.max_batch_pixels) break;
            if (keys >= s->ctx->tr_slot_cap) break;
            const int vid = it->r.volume_id;
            int a = srv_admit(s, vid);
            if (a < 0) return -a;
            if (a == 0) {
                ++it;
                continue;
            }
            if (vol_slot[vid] < 0) {
                vol_slot[vid] = (int)vols.size();
                vols.push_back(s->vols[vid].vol);
            }
            s->vols[vid].last_use = ++s->tick;
            mrp_view v = it->r.view;
            v.volume_index = vol_slot[vid];
            views.push_back(v);
            outs.push_back(it->r.dev_out4);
            px += (int64_t)v.width * v.height;
            ++keys;  // at most one TR set per request
            take.push_back(*it);
            it = s->pending.erase(it);
        }
    }
    if (take.empty()) return MRP_OK;
    int rc = mrp_prepare_views(s->ctx, vols.data(), (int)vols.size(), views.data(), (int)views.size());
    if (!rc)
        rc = mrp_render_batch(s->ctx, vols.data(), (int)vols.size(), views.data(), (int)views.size(),
                              s->cfg.precision, outs.data());
    if (rc) return rc;
    InFlight f;
    if (!s->free_events.empty()) {
        f.done = s->free_events.back();
        s->free_events.pop_back();
    } else if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess) {
        return mrp_fail(MRP_ERR_HIP, "event creation failed");
    }
    if (hipEventRecord(f.done, s->ctx->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event record failed");
    const double t = now_s();
    for (SrvReq& q : take) q.t_launch = t;
    f.reqs = std::move(take);
    s->inflight.push_back(std::move(f));
    s->stats.batches += 1;
    s->stats.batched_views += (int64_t)views.size();
    return MRP_OK;
}

// Plain mode: the drop-in per-frame FP32 launch, one request at a time in arrival order.
static int srv_launch_plain(mrp_server* s) {
    SrvReq q;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->pending.begin();
        for (; it != s->pending.end(); ++it) {
            int a = srv_admit(s, it->r.volume_id);
            if (a < 0) return -a;
            if (a == 1) break;
        }
        if (it == s->pending.end()) return MRP_OK;
        q = *it;
        s->pending.erase(it);
    }
    SrvVolume& v = s->vols[q.r.volume_id];
    v.last_use = ++s->tick;
    int rc = mrp_render_plain_device(s->ctx, v.vol, &q.r.view, q.r.dev_out4);
    if (rc) return rc;
    InFlight f;
    if (!s->free_events.empty()) {
        f.done = s->free_events.back();
        s->free_events.pop_back();
    } else if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess) {
        return mrp_fail(MRP_ERR_HIP, "event creation failed");
    }
    if (hipEventRecord(f.done, s->ctx->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "event record failed");
    q.t_launch = now_s();
    f.reqs.push_back(q);
    s->inflight.push_back(std::move(f));
    s->stats.batches += 1;
    s->stats.batched_views += 1;
    return MRP_OK;
}

extern "C" int mrp_server_step(mrp_server* s) {
    if (!s) return mrp_fail(MRP_ERR_INVALID, "server is null");
    srv_reap(s, false);
    while ((int)s->inflight.size() < s->cfg.max_inflight) {
        const size_t before = s->inflight.size();
        int rc = s->cfg.plain ? srv_launch_plain(s) : srv_launch_batch(s);
        if (rc) return rc;
        if (s->inflight.size() == before) break;  // nothing launchable
    }
    return MRP_OK;
}

extern "C" int mrp_server_drain(mrp_server* s) {
    if (!s) return mrp_fail(MRP_ERR_INVALID, "server is null");
    for (;;) {
        int rc = mrp_server_step(s);
        if (rc) return rc;
        bool empty;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            empty = s->pending.empty();
        }
        if (empty && s->inflight.empty()) return MRP_OK;
        if (!s->inflight.empty()) hipEventSynchronize(s->inflight.front().done);
        else if (!empty) {
            // pending requests whose volumes cannot be admitted even with every idle volume evicted
            bool launchable = false;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                for (const SrvReq& q : s->pending) {
                    const SrvVolume& v = s->vols[q.r.volume_id];
                    if (v.vol || v.bytes + s->cfg.hbm_reserve_bytes <= (s->cfg.hbm_limit_bytes > 0
                                                                          ? s->cfg.hbm_limit_bytes
                                                                          : INT64_MAX))
                        launchable = true;
                }
            }
            if (!launchable) return mrp_fail(MRP_ERR_NOMEM, "pending requests whose volume never fits in HBM");
        }
    }
}

extern "C" int mrp_server_poll(mrp_server* s, mrp_completion* out, int max, int* n) {
    if (!s || !n || (max > 0 && !out)) return mrp_fail(MRP_ERR_INVALID, "null argument");
    srv_reap(s, false);
    const int k = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return MRP_OK;
}

extern "C" int mrp_server_get_stats(mrp_server* s, mrp_server_stats* out) {
    if (!s || !out) return mrp_fail(MRP_ERR_INVALID, "null argument");
    *out = s->stats;
    out->resident_volumes = 0;
    for (const SrvVolume& v : s->vols) out->resident_volumes += v.vol ? 1 : 0;
    out->resident_bytes = s->resident_bytes;
    std::lock_guard<std::mutex> lk(s->mu);
    out->pending = (int64_t)s->pending.size();
    return MRP_OK;
}

extern "C" int mrp_server_is_resident(mrp_server* s, int volume_id) {
    if (!s || volume_id < 0 || volume_id >= (int)s->vols.size()) return 0;
    return s->vols[volume_id].vol ? 1 : 0;
}

// ------------------------------------------------------------------------------------------
// Open-loop load: Poisson arrivals over the server's volumes with random orbits, suns and
// resolutions, served while they arrive; per-request latency = done - arrival.
// ------------------------------------------------------------------------------------------
extern "C" int mrp_server_run_poisson(mrp_server* s, const mrp_load_spec* spec, mrp_load_report* rep) {
    if (!s || !spec || !rep) return mrp_fail(MRP_ERR_INVALID, "null argument");
    if (s->vols.empty()) return mrp_fail(MRP_ERR_STATE, "server has no volumes");
    if (spec->n_sizes < 1 || spec->n_sizes > 8 || !(spec->rate_rps > 0.0) || !(spec->duration_s > 0.0))
        return mrp_fail(MRP_ERR_INVALID, "bad load spec");
    std::mt19937_64 rng(spec->seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    // arrivals
    struct Arr {
        double t;
        mrp_request r;
        int size;
    };
    std::vector<Arr> arr;
    double t = 0.0;
    while (true) {
        t += -log(1.0 - U(rng)) / spec->rate_rps;
        if (t > spec->duration_s) break;
        Arr a;
        a.t = t;
        a.size = (int)(U(rng) * spec->n_sizes) % spec->n_sizes;
        mrp_view& v = a.r.view;
        memset(&v, 0, sizeof v);
        const double az = 2.0 * M_PI * U(rng), el = 0.6 * (U(rng) - 0.3);
        const float ori[3] = {(float)(2.0 * cos(el) * cos(az)), (float)(2.0 * sin(el)), (float)(2.0 * cos(el) * sin(az))};
        const float fl = sqrtf(ori[0] * ori[0] + ori[1] * ori[1] + ori[2] * ori[2]);
        const float fwd[3] = {-ori[0] / fl, -ori[1] / fl, -ori[2] / fl};
        float right[3] = {fwd[2], 0.0f, -fwd[0]};  // cross((0,1,0), fwd)
        const float rl = sqrtf(right[0] * right[0] + right[2] * right[2]);
        right[0] /= rl;
        right[2] /= rl;
        const float up[3] = {fwd[1] * right[2] - fwd[2] * right[1], fwd[2] * right[0] - fwd[0] * right[2],
                             fwd[0] * right[1] - fwd[1] * right[0]};
        memcpy(v.ori, ori, sizeof ori);
        memcpy(v.up, up, sizeof up);
        memcpy(v.right, right, sizeof right);
        const double sz = 0.15 + 0.8 * U(rng), sp = 2.0 * M_PI * U(rng), sr = sqrt(1.0 - sz * sz);
        v.light_dir[0] = (float)(sr * cos(sp));
        v.light_dir[1] = (float)sz;
        v.light_dir[2] = (float)(sr * sin(sp));
        v.light_color[0] = v.light_color[1] = v.light_color[2] = 1.0f;
        v.scatter_rate[0] = v.scatter_rate[1] = v.scatter_rate[2] = 0.99f;
        v.alpha = (float)(16.0 + 112.0 * U(rng));
        v.g = s->cfg.g;
        v.tan_half_fov = (float)tan(20.0 * M_PI / 180.0);
        v.jitter = 0.5f;
        v.width = spec->sizes[2 * a.size];
        v.height = spec->sizes[2 * a.size + 1];
        v.seed = (uint32_t)spec->seed;
        v.frame = 0;
        v.view_id = (uint32_t)arr.size();
        a.r.volume_id = (int)(U(rng) * s->vols.size()) % (int)s->vols.size();
        a.r.user_tag = arr.size();
        arr.push_back(a);
    }
    // output buffers: a free list per size class, allocated up front
    const int per = spec->buffers_per_size > 0 ? spec->buffers_per_size : 16;
    std::vector<std::vector<float*>> freebuf(spec->n_sizes);
    std::vector<float*> all;
    for (int k = 0; k < spec->n_sizes; ++k)
        for (int j = 0; j < per; ++j) {
            float* p = nullptr;
            if (hipMalloc(&p, (size_t)spec->sizes[2 * k] * spec->sizes[2 * k + 1] * 16) != hipSuccess) {
                for (float* q : all) hipFree(q);
                return mrp_fail(MRP_ERR_NOMEM, "output buffer allocation failed");
            }
            freebuf[k].push_back(p);
            all.push_back(p);
        }
    std::vector<float*> buf_of(arr.size(), nullptr);
    std::vector<double> lat;
    lat.reserve(arr.size());
    std::vector<mrp_completion> comp(256);
    size_t next = 0, completed = 0;
    int64_t pixels = 0;
    int rc = MRP_OK;
    const double t0 = now_s();
    while (completed < arr.size() && !rc) {
        const double el = now_s() - t0;
        // submit every request that has arrived and has an output buffer
        while (next < arr.size() && arr[next].t <= el && !freebuf[arr[next].size].empty()) {
            Arr& a = arr[next];
            a.r.dev_out4 = freebuf[a.size].back();
            freebuf[a.size].pop_back();
            buf_of[next] = a.r.dev_out4;
            if ((rc = mrp_server_submit(s, &a.r))) break;
            ++next;
        }
        if (rc) break;
        if ((rc = mrp_server_step(s))) break;
        int n = 0;
        if ((rc = mrp_server_poll(s, comp.data(), (int)comp.size(), &n))) break;
        for (int i = 0; i < n; ++i) {
            const size_t id = comp[i].user_tag;
            lat.push_back((comp[i].t_done - t0) - arr[id].t);  // from arrival, queueing included
            freebuf[arr[id].size].push_back(buf_of[id]);
            pixels += (int64_t)arr[id].r.view.width * arr[id].r.view.height;
            ++completed;
        }
        if (n == 0 && !s->inflight.empty() && (next >= arr.size() || arr[next].t > now_s() - t0))
            hipEventSynchronize(s->inflight.front().done);
    }
    const double wall = now_s() - t0;
    hipDeviceSynchronize();
    for (float* p : all) hipFree(p);
    if (rc) return rc;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(q * lat.size()))]; };
    memset(rep, 0, sizeof *rep);
    rep->requests = (int64_t)arr.size();
    rep->completed = (int64_t)completed;
    rep->pixels = pixels;
    rep->wall_s = wall;
    rep->offered_s = spec->duration_s;
    rep->mpix_s = pixels / wall / 1e6;
    rep->p50_ms = 1e3 * pct(0.50);
    rep->p99_ms = 1e3 * pct(0.99);
    rep->max_ms = lat.empty() ? 0.0 : 1e3 * lat.back();
    return MRP_OK;
}
