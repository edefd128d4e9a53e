// C ABI of libmrpnn_hip.so (declared in include/mrp_c_api.h): contexts, resident volumes,
// their preparation caches, and the render entry points.  Host code only; kernels live in
// the other translation units.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/mrp_c_api.h"
#include "capi_internal.h"
#include "fbm.h"
#include "kernels.h"
#include "net_layout.h"
#include "ref_tables.h"

using namespace mrp;

static thread_local std::string g_err;

int mrp_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return mrp_fail(MRP_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                              __FILE__, __LINE__);                                       \
    } while (0)
#define CHECK_ARG(cond, msg)                                          \
    do {                                                              \
        if (!(cond)) return mrp_fail(MRP_ERR_INVALID, "%s", (msg)); \
    } while (0)

extern "C" const char* mrp_last_error(void) { return g_err.c_str(); }
extern "C" const char* mrp_version(void) { return "mrpnn-hip 1.0 (gfx950)"; }

// ------------------------------------------------------------------------------------------
// Stencil table: GetSamples23_ (core/vector.cu:304-405) with the direction tables and the
// per-point LUT v coordinate of SetupInputData (core/radiancePredict.cu:458-464), computed
// once on the host.
// ------------------------------------------------------------------------------------------
static float lerp_h(float a, float b, float v) { return a * (1.0f - v) + b * v; }
static void build_stencil(StencilEntry* t) {
    const float ms = 0.6f;
    for (int i = 0; i < kPC; ++i) {
        StencilEntry& s = t[i];
        float off;
        s.type = 5;
        s.local = 8;
        if (i < 8) { off = (float)(1.0 / 256.0); off *= lerp_h(ms, 1.0f, 1.0f); s.mip = 0; s.local = i; }
        else if (i < 16) { off = (float)(2.5 / 256.0); off *= lerp_h(ms, 1.0f, 0.8f); s.mip = 1; }
        else if (i < 32) { off = (float)(5.5 / 256.0); off *= lerp_h(ms, 1.0f, 0.6f); s.mip = 2; }
        else if (i < 48) { off = (float)(11.5 / 256.0); off *= lerp_h(ms, 1.0f, 0.4f); s.mip = 3; }
        else if (i < 64) { off = (float)(23.5 / 256.0); off *= lerp_h(ms, 1.0f, 0.2f); s.mip = 4; }
        else if (i < 96) { off = (float)(47.5 / 256.0); off *= ms; s.mip = 5; }
        else if (i < 128) { off = (float)(95.5 / 256.0); off *= ms; s.mip = 6; }
        else if (i < 160) { off = (float)(191.5 / 256.0); off *= ms; s.mip = 7; }
        else {
            s.type = 1;
            if (i < 168) { off = (float)((i - 160) + 1); s.mip = 0; }
            else if (i < 176) { off = (float)(8.5 + 1 + 2.0 * (i - 168)); s.mip = 1; }
            else if (i < 184) { off = (float)(26.5 + 4.0 * (i - 176)); s.mip = 2; }
            else { off = (float)(60.5 + 8.0 * (i - 184)); s.mip = 3; }
            off = (float)(off / 256.0);
        }
        s.offset = off;
        if (i < 160) {
            s.dx = kStencilDirs[i][0];
            s.dy = kStencilDirs[i][1];
            s.dz = kStencilDirs[i][2];
        } else {
            s.dx = s.dy = s.dz = 0.0f;
        }
        float radius = (float)(1 << s.mip) / 256.0f;
        float angle = atanf(0.5f * radius / off);
        s.lut_v = angle / (3.1415926535f * 60.0f / 180.0f);
    }
}

// ------------------------------------------------------------------------------------------
// Weights
// ------------------------------------------------------------------------------------------
static std::string default_weights_path() {
    Dl_info info;
    if (dladdr((void*)&mrp_version, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        std::string dir = s == std::string::npos ? "." : p.substr(0, s);
        return dir + "/assets/mrpnn_weights.bin";
    }
    return "assets/mrpnn_weights.bin";
}

static int load_weights(const char* path, std::vector<float>& w) {
    std::string p = path ? path : default_weights_path();
    FILE* f = fopen(p.c_str(), "rb");
    if (!f) return mrp_fail(MRP_ERR_IO, "cannot open weight asset %s", p.c_str());
    w.resize(net::TOTAL);
    size_t got = fread(w.data(), sizeof(float), net::TOTAL, f);
    char extra;
    size_t more = fread(&extra, 1, 1, f);
    fclose(f);
    if (got != (size_t)net::TOTAL || more != 0)
        return mrp_fail(MRP_ERR_IO, "weight asset %s: expected %d floats", p.c_str(), net::TOTAL);
    uint64_t h = 0xcbf29ce484222325ull;
    const unsigned char* b = (const unsigned char*)w.data();
    for (size_t i = 0; i < w.size() * 4; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    if (h != net::FNV1A64) return mrp_fail(MRP_ERR_IO, "weight asset %s: checksum mismatch", p.c_str());
    return MRP_OK;
}

// ------------------------------------------------------------------------------------------
// Context
// ------------------------------------------------------------------------------------------
extern "C" int mrp_ctx_create(const char* weights_path, mrp_ctx** out) {
    CHECK_ARG(out, "out is null");
    *out = nullptr;
    std::vector<float> w;
    int rc = load_weights(weights_path, w);
    if (rc) return rc;
    mrp_ctx* c = new mrp_ctx();
    HIP_TRY(hipGetDevice(&c->device));
    HIP_TRY(hipGetDeviceProperties(&c->props, c->device));
    c->stream = nullptr;
    c->weights_host = w;
    if (hipMalloc(&c->d_weights, w.size() * sizeof(float)) != hipSuccess) {
        delete c;
        return mrp_fail(MRP_ERR_NOMEM, "weights allocation failed");
    }
    HIP_TRY(hipMemcpy(c->d_weights, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    StencilEntry st[kPC];
    build_stencil(st);
    HIP_TRY(hipMalloc(&c->d_stencil, sizeof st));
    HIP_TRY(hipMemcpy(c->d_stencil, st, sizeof st, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_lut, 512 * 512 * sizeof(float)));
    HIP_TRY(hipMalloc(&c->d_fast, kPC * sizeof(FastStencil)));
    HIP_TRY(hipMalloc(&c->d_lut_rows, (size_t)kPC * kLutRow * sizeof(float)));
    HIP_TRY(hipMemset(c->d_lut, 0, 512 * 512 * sizeof(float)));
    for (int i = 0; i < 4; ++i) HIP_TRY(hipEventCreate(&c->ev[i]));
    for (int i = 0; i < TimingRing::kSlots; ++i)
        for (int j = 0; j < 4; ++j) HIP_TRY(hipEventCreate(&c->ring.ev[i][j]));
    int prc = mrp_fused_init(c);
    if (prc) return prc;
    mrp_noalloc_setup(c->device);
    *out = c;
    return MRP_OK;
}

extern "C" int mrp_ctx_destroy(mrp_ctx* c) {
    if (!c) return MRP_OK;
    hipDeviceSynchronize();
    mrp_fused_free(c);
    hipFree(c->d_weights);
    hipFree(c->d_stencil);
    hipFree(c->d_lut);
    hipFree(c->d_fast);
    hipFree(c->d_lut_rows);
    hipFree(c->d_plain_out);
    hipFree(c->d_env);
    mrp_rpnn_free(c);
    for (TrSlot& t : c->tr_slots) {
        hipFree(t.d_tr);
        hipFree(t.d_ptr);
        if (t.last_read) hipEventDestroy(t.last_read);
    }
    for (int i = 0; i < 4; ++i) hipEventDestroy(c->ev[i]);
    for (int i = 0; i < TimingRing::kSlots; ++i)
        for (int j = 0; j < 4; ++j) hipEventDestroy(c->ring.ev[i][j]);
    delete c;
    return MRP_OK;
}

extern "C" int mrp_ctx_set_stream(mrp_ctx* c, void* s) {
    CHECK_ARG(c, "ctx is null");
    c->stream = (hipStream_t)s;
    return MRP_OK;
}
static int timing_read(mrp_ctx* c, int slot) {
    TimingRing& r = c->ring;
    if (!r.pending[slot]) return MRP_OK;
    HIP_TRY(hipEventSynchronize(r.ev[slot][2]));
    float a = 0.0f, b = 0.0f;
    HIP_TRY(hipEventElapsedTime(&a, r.ev[slot][0], r.ev[slot][1]));
    HIP_TRY(hipEventElapsedTime(&b, r.ev[slot][3], r.ev[slot][2]));
    r.pending[slot] = false;
    r.count += 1;
    r.sum_march += a;
    r.sum_predict += b;
    r.last_march = a;
    r.last_predict = b;
    return MRP_OK;
}
int mrp_timing_begin(mrp_ctx* c, hipEvent_t** ev) {
    TimingRing& r = c->ring;
    int rc = timing_read(c, r.head);
    if (rc) return rc;
    r.pending[r.head] = true;
    *ev = r.ev[r.head];
    r.head = (r.head + 1) % TimingRing::kSlots;
    return MRP_OK;
}
int mrp_timing_collect(mrp_ctx* c) {
    for (int k = 0; k < TimingRing::kSlots; ++k) {  // oldest first
        int rc = timing_read(c, (c->ring.head + k) % TimingRing::kSlots);
        if (rc) return rc;
    }
    c->stats.ms_march = c->ring.last_march;
    c->stats.ms_predict = c->ring.last_predict;
    c->stats.ms_total = c->ring.last_march + c->ring.last_predict;
    c->stats.timed_renders = c->ring.count;
    c->stats.ms_march_sum = c->ring.sum_march;
    c->stats.ms_predict_sum = c->ring.sum_predict;
    return MRP_OK;
}

extern "C" int mrp_ctx_set_timing(mrp_ctx* c, int enabled) {
    CHECK_ARG(c, "ctx is null");
    int rc = mrp_timing_collect(c);
    if (rc) return rc;
    c->timing = enabled != 0;
    c->ring.count = 0;
    c->ring.sum_march = c->ring.sum_predict = 0.0;
    return MRP_OK;
}
extern "C" int mrp_ctx_get_stats(mrp_ctx* c, mrp_stats* out) {
    CHECK_ARG(c && out, "null argument");
    int rc = mrp_fused_collect_stats(c);
    if (rc) return rc;
    *out = c->stats;
    return MRP_OK;
}
extern "C" int mrp_ctx_weights_device(mrp_ctx* c, const float** dev, int* count) {
    CHECK_ARG(c && dev && count, "null argument");
    *dev = c->d_weights;
    *count = net::TOTAL;
    return MRP_OK;
}

extern "C" int mrp_ctx_update_hg_lut(mrp_ctx* c, float g, uint32_t seed) {
    CHECK_ARG(c, "ctx is null");
    CHECK_ARG(g > -1.0f && g < 1.0f, "g must be in (-1, 1)");
    if (c->lut_valid && c->lut_g == g && c->lut_seed == seed) return MRP_OK;  // volume.cu:1403
    HIP_TRY(launch_hg_lut(c->d_lut, g, seed, c->stream));
    HIP_TRY(launch_fast_stencil(c->d_stencil, c->d_lut, c->d_fast, c->d_lut_rows, c->stream));
    c->lut_valid = true;
    c->lut_g = g;
    c->lut_seed = seed;
    return MRP_OK;
}
extern "C" int mrp_ctx_download_hg_lut(mrp_ctx* c, float* host) {
    CHECK_ARG(c && host, "null argument");
    HIP_TRY(hipMemcpyAsync(host, c->d_lut, 512 * 512 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MRP_OK;
}
extern "C" int mrp_ctx_upload_hg_lut(mrp_ctx* c, const float* host, float g) {
    CHECK_ARG(c && host, "null argument");
    HIP_TRY(hipMemcpyAsync(c->d_lut, host, 512 * 512 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_fast_stencil(c->d_stencil, c->d_lut, c->d_fast, c->d_lut_rows, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->lut_valid = true;
    c->lut_g = g;
    c->lut_seed = 0xffffffffu;
    return MRP_OK;
}

// ------------------------------------------------------------------------------------------
// Volumes
// ------------------------------------------------------------------------------------------

extern "C" int64_t mrp_volume_bytes(int N) {
    const int64_t on = (N + 8) >> 3;
    return (int64_t)((size_t)N * N * N + kMipFloats + kTrFloats + kPadMipFloats + kPadTrAlloc) * 4 + kQuadTexels * 16 +
           on * on * on;
}

extern "C" int mrp_volume_create(mrp_ctx* c, int N, mrp_volume** out) {
    CHECK_ARG(c && out, "null argument");
    CHECK_ARG(N >= 2 && N <= 2048, "resolution must be in [2, 2048]");
    *out = nullptr;
    mrp_volume* v = new mrp_volume();
    v->ctx = c;
    v->N = N;
    v->max_density = 0.00001f;  // volume.hpp:104
    size_t n = (size_t)N * N * N;
    if (hipMalloc(&v->d_grid, n * sizeof(float)) != hipSuccess) {
        delete v;
        return mrp_fail(MRP_ERR_NOMEM, "grid allocation (%zu bytes) failed", n * 4);
    }
    if (hipMalloc(&v->d_mips, kMipFloats * sizeof(float)) != hipSuccess ||
        hipMalloc(&v->d_tr, kTrFloats * sizeof(float)) != hipSuccess ||
        hipMalloc(&v->d_pmips, kPadMipFloats * sizeof(float)) != hipSuccess ||
        hipMalloc(&v->d_qmips, kQuadTexels * sizeof(float4)) != hipSuccess ||
        hipMalloc(&v->d_ptr, kPadTrAlloc * sizeof(float)) != hipSuccess) {
        hipFree(v->d_grid);
        hipFree(v->d_mips);
        hipFree(v->d_tr);
        hipFree(v->d_pmips);
        hipFree(v->d_qmips);
        delete v;
        return mrp_fail(MRP_ERR_NOMEM, "pyramid allocation failed");
    }
    v->occ_n = (N + 8) >> 3;
    const size_t occ_bytes = (size_t)v->occ_n * v->occ_n * v->occ_n;
    if (hipMalloc(&v->d_occ, occ_bytes) != hipSuccess) {
        hipFree(v->d_grid);
        hipFree(v->d_mips);
        hipFree(v->d_tr);
        hipFree(v->d_pmips);
        hipFree(v->d_qmips);
        hipFree(v->d_ptr);
        delete v;
        return mrp_fail(MRP_ERR_NOMEM, "occupancy allocation failed");
    }
    HIP_TRY(hipMemsetAsync(v->d_occ, 0, occ_bytes, c->stream));
    HIP_TRY(hipMemsetAsync(v->d_pmips, 0, kPadMipFloats * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(v->d_qmips, 0, kQuadTexels * sizeof(float4), c->stream));
    HIP_TRY(hipMemsetAsync(v->d_ptr, 0, kPadTrAlloc * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(v->d_grid, 0, n * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(v->d_mips, 0, kMipFloats * sizeof(float), c->stream));
    HIP_TRY(hipMemsetAsync(v->d_tr, 0, kTrFloats * sizeof(float), c->stream));
    size_t o = 0;
    for (int m = 0; m < 9; ++m) {
        v->mip[m] = v->d_mips + o;
        o += (size_t)(256 >> m) * (256 >> m) * (256 >> m);
    }
    o = 0;
    for (int m = 0; m < 8; ++m) {
        v->trm[m] = v->d_tr + o;
        o += (size_t)(128 >> m) * (128 >> m) * (128 >> m);
    }
    *out = v;
    return MRP_OK;
}

extern "C" int mrp_volume_destroy(mrp_volume* v) {
    if (!v) return MRP_OK;
    mrp_fused_join_aux(v->ctx);
    hipStreamSynchronize(v->ctx->stream);
    mrp_tr_slots_drop(v->ctx, v);
    hipFree(v->d_grid);
    hipFree(v->d_mips);
    hipFree(v->d_tr);
    hipFree(v->d_pmips);
    hipFree(v->d_qmips);
    hipFree(v->d_ptr);
    hipFree(v->d_occ);
    hipFree(v->d_ctr);
    if (v->mips_ready) hipEventDestroy(v->mips_ready);
    delete v;
    return MRP_OK;
}

static void invalidate(mrp_volume* v) {
    v->mips_valid = false;
    v->tr_valid = false;
    v->ctr_valid = false;
}

extern "C" int mrp_volume_set_data(mrp_volume* v, const float* host, float max_density) {
    CHECK_ARG(v && host, "null argument");
    CHECK_ARG(max_density > 0.0f && isfinite(max_density), "max_density must be positive and finite");
    size_t n = (size_t)v->N * v->N * v->N;
    HIP_TRY(hipMemcpyAsync(v->d_grid, host, n * sizeof(float), hipMemcpyHostToDevice, v->ctx->stream));
    HIP_TRY(hipStreamSynchronize(v->ctx->stream));
    v->max_density = max_density;
    invalidate(v);
    return MRP_OK;
}
extern "C" int mrp_volume_set_data_device(mrp_volume* v, const float* dev, float max_density) {
    CHECK_ARG(v && dev, "null argument");
    CHECK_ARG(max_density > 0.0f && isfinite(max_density), "max_density must be positive and finite");
    size_t n = (size_t)v->N * v->N * v->N;
    HIP_TRY(hipMemcpyAsync(v->d_grid, dev, n * sizeof(float), hipMemcpyDeviceToDevice, v->ctx->stream));
    v->max_density = max_density;
    invalidate(v);
    return MRP_OK;
}
extern "C" int mrp_volume_fill_fbm(mrp_volume* v, int octaves, float base_freq, uint32_t seed) {
    CHECK_ARG(v, "volume is null");
    CHECK_ARG(octaves >= 1 && octaves <= 16, "octaves must be in [1, 16]");
    int perm[512];
    fbm_make_perm(seed, perm);
    int* d_perm = nullptr;
    uint32_t* d_mm = nullptr;
    HIP_TRY(hipMalloc(&d_perm, sizeof perm));
    HIP_TRY(hipMalloc(&d_mm, 2 * sizeof(uint32_t)));
    uint32_t init[2] = {0xffffffffu, 0u};
    hipStream_t s = v->ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_perm, perm, sizeof perm, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_mm, init, sizeof init, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_fbm(v->d_grid, v->N, octaves, base_freq, d_perm, d_mm, s));
    HIP_TRY(hipStreamSynchronize(s));
    hipFree(d_perm);
    hipFree(d_mm);
    v->max_density = 1.0f;
    invalidate(v);
    return MRP_OK;
}
extern "C" int mrp_volume_get_data(mrp_volume* v, float* host) {
    CHECK_ARG(v && host, "null argument");
    size_t n = (size_t)v->N * v->N * v->N;
    HIP_TRY(hipMemcpyAsync(host, v->d_grid, n * sizeof(float), hipMemcpyDeviceToHost, v->ctx->stream));
    HIP_TRY(hipStreamSynchronize(v->ctx->stream));
    return MRP_OK;
}
extern "C" int mrp_volume_resolution(mrp_volume* v) { return v ? v->N : -1; }
extern "C" const float* mrp_volume_grid_device(mrp_volume* v) { return v ? v->d_grid : nullptr; }

// Update(process_mips) (core/volume.cu:1443-1531)
extern "C" int mrp_volume_update(mrp_volume* v) {
    CHECK_ARG(v, "volume is null");
    hipStream_t s = v->ctx->stream;
    int jrc = mrp_fused_join_aux(v->ctx);  // TR bakes in flight read the old pyramid
    if (jrc) return jrc;
    int res = 256;
    while (res * 2 < v->N) res <<= 1;
    const float* src = v->d_grid;
    int sres = v->N;
    float* tmp[2] = {nullptr, nullptr};
    int which = 0;
    for (; res > 256; res >>= 1) {  // intermediate power-of-two levels above 256 (volume.cu:1476-1493)
        float*& t = tmp[which];
        if (!t) HIP_TRY(hipMalloc(&t, (size_t)res * res * res * sizeof(float)));
        if (sres == 2 * res) HIP_TRY(launch_box2(src, t, res, s));
        else HIP_TRY(launch_resample(src, sres, t, res, s));
        src = t;
        sres = res;
        which ^= 1;
    }
    for (int m = 0; m < 9; ++m) {
        int r = 256 >> m;
        if (sres == 2 * r) HIP_TRY(launch_box2(src, v->mip[m], r, s));
        else HIP_TRY(launch_resample(src, sres, v->mip[m], r, s));
        src = v->mip[m];
        sres = r;
    }
    {
        const float* mips[9];
        for (int m = 0; m < 9; ++m) mips[m] = v->mip[m];
        HIP_TRY(launch_pad_mips(mips, v->d_pmips, s));
        HIP_TRY(launch_quad_mips(v->d_pmips, v->d_qmips, s));
    }
    HIP_TRY(launch_occupancy(v->d_grid, v->N, v->d_occ, v->occ_n, s));
    if (tmp[0] || tmp[1]) {
        HIP_TRY(hipStreamSynchronize(s));
        hipFree(tmp[0]);
        hipFree(tmp[1]);
    }
    if (!v->mips_ready) HIP_TRY(hipEventCreateWithFlags(&v->mips_ready, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(v->mips_ready, s));
    v->mips_valid = true;
    ++v->mips_gen;
    v->tr_valid = false;
    v->ctr_valid = false;
    mrp_tr_slots_drop(v->ctx, v);
    return MRP_OK;
}

extern "C" int mrp_volume_get_info(const mrp_volume* v, mrp_volume_info* out) {
    CHECK_ARG(v && out, "null argument");
    memset(out, 0, sizeof *out);
    out->resolution = v->N;
    out->max_density = v->max_density;
    out->mips_valid = v->mips_valid ? 1 : 0;
    out->mips_generation = v->mips_gen;
    out->tr_valid = v->tr_valid ? 1 : 0;
    memcpy(out->tr_light_dir, v->tr_light, sizeof out->tr_light_dir);
    out->tr_alpha = v->tr_alpha;
    out->cached_tr_res = v->ctr_valid ? v->ctr_res : 0;
    return MRP_OK;
}

extern "C" int mrp_ctx_get_hg_lut_g(mrp_ctx* c, float* g) {
    CHECK_ARG(c && g, "null argument");
    if (!c->lut_valid) return mrp_fail(MRP_ERR_STATE, "no HG LUT built");
    *g = c->lut_g;
    return MRP_OK;
}

// Update_TR (core/volume.cu:1589-1667), cached on exact (light, alpha)
// Fill_TR for all 8 levels into the given buffers (levels packed as in mrp_volume::d_tr).
int mrp_tr_bake_into(mrp_volume* v, const float L[3], float alpha, float* d_tr, float* d_ptr, hipStream_t s) {
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before the TR bake");
    float alpha_lvl[8];
    for (int m = 0; m < 8; ++m) alpha_lvl[m] = alpha / powf(1.73f, (float)m + 1.0f);
    const float* mips[9];
    float* tr[8];
    for (int m = 0; m < 9; ++m) mips[m] = v->mip[m];
    size_t o = 0;
    for (int m = 0; m < 8; ++m) {
        tr[m] = d_tr + o;
        o += (size_t)(128 >> m) * (128 >> m) * (128 >> m);
    }
    HIP_TRY(launch_tr_bake(mips, v->d_pmips, v->d_qmips, tr, d_ptr, alpha_lvl, mk3(L[0], L[1], L[2]), s));
    return MRP_OK;
}

int mrp_volume_update_tr_impl(mrp_volume* v, const float L[3], float alpha) {
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before the TR bake");
    if (v->tr_valid && v->tr_light[0] == L[0] && v->tr_light[1] == L[1] && v->tr_light[2] == L[2] &&
        v->tr_alpha == alpha)
        return MRP_OK;
    int rc = mrp_tr_bake_into(v, L, alpha, v->d_tr, v->d_ptr, v->ctx->stream);
    if (rc) return rc;
    v->tr_valid = true;
    memcpy(v->tr_light, L, sizeof v->tr_light);
    v->tr_alpha = alpha;
    return MRP_OK;
}
extern "C" int mrp_volume_update_tr(mrp_volume* v, const float L[3], float alpha) {
    CHECK_ARG(v && L, "null argument");
    CHECK_ARG(alpha > 0.0f && isfinite(alpha), "alpha must be positive and finite");
    return mrp_volume_update_tr_impl(v, L, alpha);
}
extern "C" int mrp_volume_download_mips(mrp_volume* v, float* host) {
    CHECK_ARG(v && host, "null argument");
    HIP_TRY(hipMemcpyAsync(host, v->d_mips, kMipFloats * 4, hipMemcpyDeviceToHost, v->ctx->stream));
    HIP_TRY(hipStreamSynchronize(v->ctx->stream));
    return MRP_OK;
}
extern "C" int mrp_volume_download_tr(mrp_volume* v, float* host) {
    CHECK_ARG(v && host, "null argument");
    HIP_TRY(hipMemcpyAsync(host, v->d_tr, kTrFloats * 4, hipMemcpyDeviceToHost, v->ctx->stream));
    HIP_TRY(hipStreamSynchronize(v->ctx->stream));
    return MRP_OK;
}

// ------------------------------------------------------------------------------------------
// Views
// ------------------------------------------------------------------------------------------
int mrp_check_view(const mrp_view* vw) {
    CHECK_ARG(vw, "view is null");
    CHECK_ARG(vw->width >= 1 && vw->height >= 1 && vw->width <= 16384 && vw->height <= 16384,
              "image size must be in [1, 16384]");
    CHECK_ARG(vw->alpha > 0.0f && isfinite(vw->alpha), "alpha must be positive and finite");
    CHECK_ARG(vw->g > -1.0f && vw->g < 1.0f, "g must be in (-1, 1)");
    CHECK_ARG(vw->tan_half_fov > 0.0f && isfinite(vw->tan_half_fov), "tan_half_fov must be positive");
    float l2 = vw->light_dir[0] * vw->light_dir[0] + vw->light_dir[1] * vw->light_dir[1] +
               vw->light_dir[2] * vw->light_dir[2];
    CHECK_ARG(l2 > 0.0f && isfinite(l2), "light_dir must be non-zero");
    return MRP_OK;
}
ViewDev mrp_to_dev(const mrp_view* v) {
    ViewDev d;
    d.ori = mk3(v->ori[0], v->ori[1], v->ori[2]);
    d.up = mk3(v->up[0], v->up[1], v->up[2]);
    d.right = mk3(v->right[0], v->right[1], v->right[2]);
    d.light_dir = mk3(v->light_dir[0], v->light_dir[1], v->light_dir[2]);
    d.light_color = mk3(v->light_color[0], v->light_color[1], v->light_color[2]);
    d.scatter_rate = mk3(v->scatter_rate[0], v->scatter_rate[1], v->scatter_rate[2]);
    d.alpha = v->alpha;
    d.g = v->g;
    d.tan_half_fov = v->tan_half_fov;
    d.jitter = v->jitter;
    d.width = v->width;
    d.height = v->height;
    d.seed = v->seed;
    d.frame = v->frame;
    d.view_id = v->view_id;
    d.volume_index = v->volume_index;
    return d;
}
DevVolume mrp_dev_volume(const mrp_volume* v) {
    DevVolume d;
    d.grid = v->d_grid;
    for (int m = 0; m < 9; ++m) d.mips[m] = v->mip[m];
    for (int m = 0; m < 8; ++m) d.tr[m] = v->trm[m];
    d.lut = v->ctx->d_lut;
    d.ptr = v->d_ptr;
    d.qmips = v->d_qmips;
    d.occ = v->d_occ;
    d.occ_n = v->occ_n;
    d.N = v->N;
    d.max_density = v->max_density;
    d.env = v->ctx->d_env;
    d.env_w = v->ctx->env_w;
    d.env_h = v->ctx->env_h;
    d.env_exp = v->ctx->env_exp;
    return d;
}

// LUT for g and TR for (light, alpha) of one view (VolumeRender::Render, volume.cu:511-514)
int mrp_prepare_view(mrp_ctx* c, mrp_volume* v, const mrp_view* vw) {
    if (!v->mips_valid) return mrp_fail(MRP_ERR_STATE, "mrp_volume_update must run before rendering");
    int rc = mrp_ctx_update_hg_lut(c, vw->g, c->lut_valid ? c->lut_seed : 1u);
    if (rc) return rc;
    return mrp_volume_update_tr_impl(v, vw->light_dir, vw->alpha);
}

// ------------------------------------------------------------------------------------------
// Plain per-frame FP32 path (the drop-in): VolumeRender::Render (core/volume.cu:487-548)
// ------------------------------------------------------------------------------------------
extern "C" int mrp_render_plain_device(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, float* dev_out4) {
    CHECK_ARG(c && v && dev_out4, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    // LUT for g and TR mips for (light, alpha), rebuilt only when they change (volume.cu:511-514)
    rc = mrp_prepare_view(c, v, vw);
    if (rc) return rc;
    DevVolume dv = mrp_dev_volume(v);
    ViewDev d = mrp_to_dev(vw);
    hipEvent_t* ev = nullptr;
    if (c->timing) {
        rc = mrp_timing_begin(c, &ev);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ev[0], c->stream));
        HIP_TRY(hipEventRecord(ev[1], c->stream));
        HIP_TRY(hipEventRecord(ev[3], c->stream));
    }
    HIP_TRY(launch_predict_plain(dv, c->d_stencil, c->d_weights, d, dev_out4, 1.0f, 1, c->stream));
    if (c->timing) HIP_TRY(hipEventRecord(ev[2], c->stream));
    c->stats.pixels = (int64_t)vw->width * vw->height;
    c->stats.active_pixels = -1;
    c->stats.march_taps = -1;
    return MRP_OK;
}

extern "C" int mrp_render_plain(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, int n_frames, float* host_out4) {
    CHECK_ARG(c && v && host_out4, "null argument");
    CHECK_ARG(n_frames >= 1, "n_frames must be >= 1");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    size_t bytes = (size_t)vw->width * vw->height * 4 * sizeof(float);
    if (bytes > c->plain_out_bytes) {
        hipFree(c->d_plain_out);
        c->d_plain_out = nullptr;
        if (hipMalloc(&c->d_plain_out, bytes) != hipSuccess) return mrp_fail(MRP_ERR_NOMEM, "output allocation failed");
        c->plain_out_bytes = bytes;
    }
    DevVolume dv = mrp_dev_volume(v);
    for (int f = 0; f < n_frames; ++f) {
        rc = mrp_prepare_view(c, v, vw);
        if (rc) return rc;
        mrp_view vf = *vw;
        vf.frame = vw->frame + (uint32_t)f;
        ViewDev d = mrp_to_dev(&vf);
        HIP_TRY(launch_predict_plain(dv, c->d_stencil, c->d_weights, d, c->d_plain_out, 1.0f / (1.0f + (float)f),
                                     f == 0, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // blocking per frame (volume.cu:536)
    }
    HIP_TRY(hipMemcpyAsync(host_out4, c->d_plain_out, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->stats.pixels = (int64_t)vw->width * vw->height * n_frames;
    c->stats.active_pixels = -1;
    c->stats.march_taps = -1;
    return MRP_OK;
}

extern "C" int mrp_gather_features(mrp_ctx* c, mrp_volume* v, const mrp_view* vw, float* dev_feats, int32_t* dev_active,
                                   float* dev_pos3, int32_t* dev_taps) {
    CHECK_ARG(c && v && dev_feats, "null argument");
    int rc = mrp_check_view(vw);
    if (rc) return rc;
    if (!v->mips_valid || !c->lut_valid || !v->tr_valid)
        return mrp_fail(MRP_ERR_STATE, "LUT / TR mips not prepared for this view");
    HIP_TRY(launch_gather_features(mrp_dev_volume(v), c->d_stencil, mrp_to_dev(vw), dev_feats, dev_active, dev_pos3,
                                   dev_taps, c->stream));
    return MRP_OK;
}

extern "C" int mrp_predict_features(mrp_ctx* c, const float* dev_feats, int n, const float sr[3], float g,
                                    int precision, float* dev_rgb) {
    CHECK_ARG(c && dev_feats && dev_rgb && sr, "null argument");
    CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return MRP_OK;
    if (precision == MRP_FP32) {
        HIP_TRY(launch_predict_features_fp32(dev_feats, n, c->d_weights, mk3(sr[0], sr[1], sr[2]), g, dev_rgb,
                                             c->stream));
        return MRP_OK;
    }
    return mrp_fused_predict_features(c, dev_feats, n, sr, g, precision, dev_rgb);
}
