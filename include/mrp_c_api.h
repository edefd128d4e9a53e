/* C ABI of the MI355X (gfx950) MRPNN radiance-prediction library, libmrpnn_hip.so.
 *
 * This is the boundary the reference's C++ classes VolumeRender / Camera
 * (core/volume.hpp:45-155, core/camera.hpp:10-43) are rebuilt on (see include/mrpnn/).
 * Plain C types only: opaque handles, int status codes, caller-owned buffers.
 *
 * Conventions
 *   - Every function returns MRP_OK (0) or an MRP_ERR_* code; mrp_last_error() gives the
 *     message of the last failure on the calling thread.  Nothing prints and carries on
 *     (the reference's CheckError does, core/volume.cu:15).
 *   - Volume grids are float32 N^3 in the reference layout: index (x*N + y)*N + z,
 *     x slowest (core/volume.cu:1179).  Position p in [-0.5,0.5]^3 maps to x <- p.x.
 *   - "dev_" pointers are HIP device pointers; all device work is enqueued on the context's
 *     stream (mrp_ctx_set_stream) and the *_device calls return without synchronising.
 *     Host-buffer calls synchronise before returning.
 *   - A context is not thread-safe; use one per thread.  Volumes belong to one context.
 */
#ifndef MRP_C_API_H
#define MRP_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRP_OK 0
#define MRP_ERR_INVALID 1 /* bad argument (null pointer, size out of range, ...) */
#define MRP_ERR_HIP 2     /* HIP runtime error */
#define MRP_ERR_NOMEM 3   /* device allocation failed */
#define MRP_ERR_IO 4      /* weight asset missing or malformed */
#define MRP_ERR_STATE 5   /* call out of order, e.g. render before mrp_volume_update */

/* MLP arithmetic of the fused predictor. */
#define MRP_FP32 0    /* scalar fp32, the reference's arithmetic */
#define MRP_BF16 1    /* bf16 MFMA, fp32 accumulation; the tail layers after LC0 (LC1, LC2, LX, LX0..LX12)
                         with weights as two bf16 terms (hi + lo) */
#define MRP_FP8 2     /* fp8 e4m3 MFMA, per-layer max-abs scales of weights and of activations
                         (calibrated, see mrp_fp8_calibrate), fp32 accumulation */
#define MRP_BF16_1T 3 /* bf16 MFMA with single-term bf16 weights (fastest, least accurate) */

/* Context options (mrp_ctx_set_option).  No counterpart in the reference, whose launch
 * shapes are fixed at compile time. */
#define MRP_OPT_MFMA_WAVES_PER_CU 1 /* persistent-grid waves per CU of the fused MFMA kernel (1..32, default 12) */
#define MRP_OPT_EXACT_GATHER 2      /* 1: the fused MFMA kernels gather features with the FP32 path's exact
                                       arithmetic (bit-identical features); 0 (default): fast gather from
                                       zero-padded pyramid copies (features within a few ulps) */
#define MRP_OPT_GRAPH 4             /* 1: mrp_render_batch captures its work (staging copies, counter reset,
                                       march, network) into a HIP graph once per launch shape and
                                       replays it; timing then reports the whole replay as ms_predict */
#define MRP_OPT_TR_SLOTS 3          /* capacity of the batched path's TR cache: one set of transmittance
                                       mips and their quads (about 51 MB) per distinct (volume, light_dir, alpha) of a
                                       batch (1..4096, default 64) */
#define MRP_OPT_LDS_WEIGHTS 5       /* 1 (default): the fused MFMA kernel runs one 768-thread workgroup
                                       per CU holding all weight fragments and biases in LDS;
                                       0: fragments are read through L1/L2 by MRP_OPT_MFMA_WAVES_PER_CU
                                       persistent waves per CU */
#define MRP_OPT_TILE_W 7            /* width in pixels of the batched path's march tiles, one wavefront of
                                       64 pixels each: 4, 8 (default), 16, 32 or 64 (height 64 / width) */
#define MRP_OPT_RPNN_CHUNK 6        /* pixels per pass of the BF16 RPNN path (descriptor fragments of
                                       5 KB per pixel, then the network): 4096..1048576, a multiple
                                       of 128 (default 65536) */

/* Number of features per pixel: [X_Val 0:192 | X_Val_Hg 192:384 | X_Val_Sub 384:576 | gamma]
 * (core/radiancePredict.cu:427-485, row order of PopulateInputData :562-565). */
#define MRP_NFEAT 577

typedef struct mrp_ctx mrp_ctx;
typedef struct mrp_volume mrp_volume;

/* One camera view and its lighting; the argument list of VolumeRender::Render
 * (core/volume.hpp:151) plus the device globals it reads (scatter_rate, core/render.cuh:17).
 * Ray of pixel (row, col):  u = 1-(col+jx)/width, v = 1-(row+jy)/height,
 *   dir = normalize(normalize(-ori) + right*((2u-1)*tan_half_fov*width/height)
 *                                   + up*((2v-1)*tan_half_fov))
 * tan_half_fov = 1 and width == height reproduce the reference camera
 * (core/volume.cu:206-216, 90 degree square).  jitter < 0 draws (jx,jy) from the pixel's
 * Philox stream, otherwise jx = jy = jitter (0.5 = pixel centre).  Output pixel index is
 * row*width + col. */
typedef struct mrp_view {
    float ori[3];
    float up[3];
    float right[3];
    float light_dir[3];
    float light_color[3];
    float scatter_rate[3]; /* albedo per channel, SetScatterRate (core/volume.cu:1748-1756) */
    float alpha;           /* extinction scale sigma_t / max_density */
    float g;               /* Henyey-Greenstein anisotropy */
    float tan_half_fov;
    float jitter;
    int32_t width;
    int32_t height;
    uint32_t seed;         /* Philox key word 0 */
    uint32_t frame;        /* Philox key word 1 (the reference's frame counter) */
    uint32_t view_id;      /* Philox counter word; keeps a view's result independent of batching */
    int32_t volume_index;  /* index into the volume array of mrp_render_batch */
} mrp_view;

/* Statistics of the last render on a context (counts, device times) and device-time sums over
 * the renders timed since timing was switched on.  Timing records HIP events on the
 * context's stream and never blocks a render; mrp_ctx_get_stats waits for them. */
typedef struct mrp_stats {
    int64_t pixels;        /* width*height summed over views */
    int64_t active_pixels; /* pixels whose ray reached a first-scattering point */
    int64_t march_taps;    /* full-resolution density taps of the first-hit march (S_march) */
    float ms_march;        /* device time of the march kernel (when timing is on) */
    float ms_predict;      /* device time of the gather+MLP kernel(s) */
    float ms_total;        /* device time of the whole render on the stream */
    int64_t timed_renders; /* renders timed since timing was last switched (mrp_ctx_set_timing) */
    double ms_march_sum;   /* their summed march time */
    double ms_predict_sum; /* their summed gather+MLP time */
} mrp_stats;

const char* mrp_last_error(void);
const char* mrp_version(void);

/* Context: device weights, HG LUT, stream, scratch.  weights_path NULL = the asset
 * mrpnn_weights.bin installed in assets/ next to the library. */
int mrp_ctx_create(const char* weights_path, mrp_ctx** out);
/* Volumes and servers of a context must be destroyed before the context. */
int mrp_ctx_destroy(mrp_ctx* ctx);
int mrp_ctx_set_stream(mrp_ctx* ctx, void* hip_stream);
int mrp_ctx_set_timing(mrp_ctx* ctx, int enabled);
int mrp_ctx_get_stats(mrp_ctx* ctx, mrp_stats* out);
int mrp_ctx_set_option(mrp_ctx* ctx, int option, int value);
int mrp_ctx_get_option(mrp_ctx* ctx, int option, int* value);
/* Device pointer of the MLP weights as loaded (fp32, header order). */
int mrp_ctx_weights_device(mrp_ctx* ctx, const float** dev_weights, int* count);

/* HG LUT, part of VolumeRender::UpdateHGLut / Fill_Hg (core/volume.cu:1322-1423): 512x512
 * mean HG phase over a ball, cached on exact g.  seed keys the per-texel jitter. */
int mrp_ctx_update_hg_lut(mrp_ctx* ctx, float g, uint32_t seed);
int mrp_ctx_download_hg_lut(mrp_ctx* ctx, float* host_lut /* 512*512 */);
int mrp_ctx_upload_hg_lut(mrp_ctx* ctx, const float* host_lut, float g); /* testing hook */

/* Volumes: VolumeRender(int resolution) (core/volume.cu:948-957) without its 1024^3
 * scratch array.  N must be in [2, 2048]. */
int mrp_volume_create(mrp_ctx* ctx, int resolution, mrp_volume** out);
int mrp_volume_destroy(mrp_volume* vol);
/* Device bytes mrp_volume_create(N) allocates (grid + mips + TR mips). */
int64_t mrp_volume_bytes(int resolution);
/* SetDatas/SetData + max_density field (core/volume.cu:1178-1184, volume.hpp:104). */
int mrp_volume_set_data(mrp_volume* vol, const float* host_grid, float max_density);
int mrp_volume_set_data_device(mrp_volume* vol, const float* dev_grid, float max_density);
/* Seeded fBm cloud generated on the device (octaves, base frequency, seed); sets
 * max_density = 1.  Same bits as the CPU generator of the test oracle. */
int mrp_volume_fill_fbm(mrp_volume* vol, int octaves, float base_freq, uint32_t seed);
int mrp_volume_get_data(mrp_volume* vol, float* host_grid);
/* Update(process_mips) (core/volume.cu:1443-1531): mip pyramid 256^3..1^3 on the GPU. */
int mrp_volume_update(mrp_volume* vol);
/* Update_TR (core/volume.cu:1589-1667): 8 transmittance mips for (light_dir, alpha),
 * cached on exact equality like the reference. */
int mrp_volume_update_tr(mrp_volume* vol, const float light_dir[3], float alpha);
/* Copies of the device pyramids: mips concatenated 256^3,128^3,...,1^3 (19,173,961 floats);
 * TR mips 128^3..1^3 (2,396,745 floats). */
int mrp_volume_download_mips(mrp_volume* vol, float* host_mips);
int mrp_volume_download_tr(mrp_volume* vol, float* host_tr);
int mrp_volume_resolution(mrp_volume* vol);
const float* mrp_volume_grid_device(mrp_volume* vol);
/* State of a volume's device pyramids (no counterpart: the reference keeps host copies of its
 * mips / TR mips next to the device ones and reads them directly).  mips_generation counts
 * successful mrp_volume_update calls, so a host copy can tell when it went stale. */
typedef struct mrp_volume_info {
    int32_t resolution;
    float max_density;
    int32_t mips_valid;
    uint32_t mips_generation;
    int32_t tr_valid;           /* TR mips of mrp_volume_update_tr present for ... */
    float tr_light_dir[3];      /* ... this light_dir and */
    float tr_alpha;             /* ... this alpha */
    int32_t cached_tr_res;      /* mrp_volume_bake_cached_tr grid, 0 when none */
} mrp_volume_info;
int mrp_volume_get_info(const mrp_volume* vol, mrp_volume_info* out);
/* g of the context's current HG LUT; MRP_ERR_STATE when none was built. */
int mrp_ctx_get_hg_lut_g(mrp_ctx* ctx, float* g);

/* Plain per-frame FP32 predictor: VolumeRender::Render(size, ..., MRPNN, sampleNum)
 * (core/volume.cu:487-548): for each frame, rebuild LUT/TR if (g, light, alpha) changed,
 * one thread per pixel for ray, march, basis, gather and the scalar fp32 MLP, progressive
 * average over frames (frame f uses Philox key word 1 = view->frame + f).  Blocking;
 * host_out4 receives width*height float4 (rgb, distance). */
int mrp_render_plain(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, int n_frames, float* host_out4);
/* One frame of the same kernel into a device buffer, asynchronous; rebuilds the LUT and the
 * volume's TR mips first when (g, light, alpha) changed, as mrp_render_plain does. */
int mrp_render_plain_device(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, float* dev_out4);

/* Fused predictor: march + compaction, then gather and MLP (FP32 / BF16 MFMA / FP8 MFMA)
 * for all views in one pass; views[i].volume_index selects volumes[...].  dev_out4[i] has
 * width*height float4.  LUT and TR mips must be current for each view (see
 * mrp_prepare_views).  Asynchronous. */
int mrp_render_batch(mrp_ctx* ctx, mrp_volume* const* volumes, int n_volumes, const mrp_view* views, int n_views,
                     int precision, float* const* dev_out4);
/* mrp_render_batch without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels, arguments and results as mrp_render_batch, but the
 * per-call device buffers (counters, volume and view tables, the active-pixel list of 64-byte
 * records, 64 per march tile) live in one static device array of MRP_NOALLOC_SCRATCH_BYTES that
 * the library keeps on each GPU it runs on until the process ends (resolved by
 * mrp_ctx_create).  Limits, refused with MRP_ERR_INVALID: n_views <= MRP_NOALLOC_MAX_VIEWS and
 * sum over views of 64 * (march tiles of the view) <= MRP_NOALLOC_MAX_RECORDS (one 1920x1080
 * view needs 2,073,600 with 8x8 tiles, 2,088,960 with 4x16 tiles).  The views must be prepared (mrp_prepare_views) as for
 * mrp_render_batch; MRP_OPT_GRAPH and timing are ignored.  The array is shared by every call on
 * one GPU and keeps the last call's contents: calls that overlap on one GPU (other streams or
 * contexts) race; calls on one stream are ordered.  Asynchronous. */
#define MRP_NOALLOC_MAX_VIEWS 64
#define MRP_NOALLOC_MAX_RECORDS 2097152
#define MRP_NOALLOC_SCRATCH_BYTES (65536 + (size_t)MRP_NOALLOC_MAX_RECORDS * 64)
int mrp_render_batch_noalloc(mrp_ctx* ctx, mrp_volume* const* volumes, int n_volumes, const mrp_view* views,
                             int n_views, int precision, float* const* dev_out4);
/* Launch-shape autotuner (no counterpart in the reference, whose launch shapes are fixed):
 * renders the prepared batch first with march tiles 8x8, 16x4, 4x16, 32x2 and 64x1 and keeps the
 * fastest (MRP_OPT_TILE_W), then with the LDS-weight network kernel (12 waves per CU) and the
 * L1/L2-weight kernel at 4, 8, 12, 16, 24 and 32 persistent waves per CU and keeps the fastest
 * (MRP_OPT_LDS_WEIGHTS, MRP_OPT_MFMA_WAVES_PER_CU); reports the network's waves per CU and the
 * time per batch of the chosen shape.  Images do not depend on the shape.  Blocking. */
int mrp_autotune_batch(mrp_ctx* ctx, mrp_volume* const* volumes, int n_volumes, const mrp_view* views, int n_views,
                       int precision, float* const* dev_out4, int* best_waves_per_cu, float* best_ms);
/* Brings the HG LUT (all views share g) and the batch TR cache up to date: one set of TR mips
 * per distinct (volume, light_dir, alpha) of the views, baked only when not cached
 * (Update_TR, core/volume.cu:1589-1667).  Views of one volume may use different suns.  Sets
 * least recently used by earlier calls are reused when the cache (MRP_OPT_TR_SLOTS) is full;
 * MRP_ERR_NOMEM when one call needs more sets than the cache holds. */
int mrp_prepare_views(mrp_ctx* ctx, mrp_volume* const* volumes, int n_volumes, const mrp_view* views, int n_views);

/* Feature-level entry points (parity and tooling).
 * mrp_gather_features: per pixel the 577 features (zeros for inactive pixels),
 *   active flag, hit position, and march taps, exactly as the plain kernel computes them.
 * mrp_predict_features: RadiancePredictInternal on n feature rows (577 floats each) with
 *   the given scatterrate (the value the reference passes: scatter_rate/1.001) and g;
 *   dev_rgb receives n*3 floats max(e^P-1,0)*scatterrate^4. */
int mrp_gather_features(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, float* dev_feats, int32_t* dev_active,
                        float* dev_pos3, int32_t* dev_taps);
int mrp_predict_features(mrp_ctx* ctx, const float* dev_feats, int n, const float scatterrate[3], float g, int precision,
                         float* dev_rgb);
/* FP8 activation scales (no counterpart in the reference, whose networks run in FP32).  The
 * MRP_FP8 network stores the inputs of each of its MRP_FP8_LAYERS MFMA layers (the 27 trunk
 * layers L01..L81 of the three branches, the 12 DI layers, LD41 / LD_Tr41 / LD_Hg41, LSEFin2,
 * LC0, LC1, LC2, LX, LX0..LX11, LX12, in that order) as e4m3 values of clamp(x, +-448*2^-e) * 2^e
 * with one exponent e per layer: max-abs scaling, where the largest |input| the layer saw in
 * calibration maps to at most 448, e4m3's largest finite value.  A context starts with the
 * exponents calibrated on seeded fBm clouds that ship with the library.
 * mrp_fp8_calibrate: runs the BF16 network over n feature rows (577 floats each, the
 *   mrp_predict_features format) recording per layer the largest |input|; reset != 0 clears the
 *   recorded maxima first.  Installs the exponents of all maxima recorded since the last reset
 *   (the largest e with max * 2^e <= 448) and, when maxima is not null, writes the
 *   MRP_FP8_LAYERS maxima there.  Blocking; it orders itself after earlier work on the context's
 *   stream, and work issued on other streams that reads the exponents must not overlap it.
 * mrp_fp8_get_exponents / mrp_fp8_set_exponents: the MRP_FP8_LAYERS exponents (n must be
 *   MRP_FP8_LAYERS; each in [-100, 100]).  Setting is blocking like calibration. */
#define MRP_FP8_LAYERS 60
int mrp_fp8_calibrate(mrp_ctx* ctx, const float* dev_feats, int n, const float scatterrate[3], float g, int reset,
                      float* maxima);
int mrp_fp8_get_exponents(mrp_ctx* ctx, int* exps, int n);
int mrp_fp8_set_exponents(mrp_ctx* ctx, const int* exps, int n);

/* Tooling (no counterpart in the reference): the 577 features of each pixel as the fused MFMA
 * kernels gather them by default (fast gather: zero-padded pyramid copies, FMA lerps, hardware
 * log), for feature-by-feature checks against mrp_gather_features.  Prepares the view's LUT and
 * TR set (mrp_prepare_views) and runs one BF16 batch for its active-pixel list; rows of pixels
 * without a first-scattering point are not written.  Blocking. */
int mrp_gather_features_fast(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, float* dev_feats);

/* Tile feature gather, the training-data path: VolumeRender::GatherData (core/volume.cu:630-708)
 * -> GatherSamples (core/volume.cu:257-283) -> NNGatherSamples / PrepareRay / GetSampleT
 * (core/render.cuh:1006-1142, 464-486, 199-397) -> SetupInputData2 (core/radiancePredict.cu:487-565),
 * called by the reference's MRPNNInputData.fetch_input (mrpnn_input_data_wrapper.cpp:72-84).
 * Brings the LUT (view->g) and the volume's TR mips (light_dir, alpha) up to date, then writes
 * width*height rows of MRP_TILE_ROW floats to dev_out: [X_Val 0:192 | X_Val_Hg 192:384 |
 * X_Val_Sub 384:576 | gamma, d_gamma_right, d_gamma_up, horizontal t diff, vertical t diff |
 * per sub-ray of the 8x8 grid (mid t, t range) 581:709 | per sub-ray opacity 709:773]
 * (mrpnn_helpers.py unpack_features).  pixel_offset_x/y shift the pixel grid by eighths of a
 * pixel.  A pixel none of whose 64 sub-rays hits the medium gets row[0] = -1000, zeros in
 * 1:581 (the reference leaves them unwritten) and its t grid.  Asynchronous. */
#define MRP_TILE_ROW 773
int mrp_gather_tiles(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, int pixel_offset_x, int pixel_offset_y,
                     float* dev_out);

/* ---------------------------------------------------------------------------------------
 * RPNN, the paper's comparison predictor: RenderCamera<true, RPNN> -> NNPredict<RPNN> ->
 * RadiancePredict_RPNN (core/volume.cu:198-228, core/render.cuh:808-875,
 * core/radiancePredict.cu:567-722), weights of core/NNWeight_RPNN.cuh as the asset
 * rpnn_weights.bin (1,296,401 floats, FNV-1a checked).
 * --------------------------------------------------------------------------------------- */
/* Loads the asset (path NULL: assets/rpnn_weights.bin next to the library); the render calls
 * load it on first use. */
int mrp_ctx_load_rpnn(mrp_ctx* ctx, const char* path);
/* One frame into dev_out4 (width*height float4: max(0, radiance * light_color * scatter_rate),
 * distance; sky = SkyBox).  precision MRP_FP32: the reference's arithmetic (every output sums
 * its products from input 0 upwards, bias last); MRP_BF16: bf16 MFMA, fp32 accumulation.
 * Synchronises with the march (the active-pixel count sizes the network launches). */
int mrp_render_rpnn(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, int precision, float* dev_out4);
/* Per pixel the network's input row of 2251 floats (10 levels x 225 density samples, gamma;
 * zeros for sky pixels) into dev_rows, and the sky pixels of dev_out4.  Parity tooling. */
int mrp_rpnn_features(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, float* dev_rows, float* dev_out4);
/* The network alone on n given rows of 2251 floats: dev_y[i] = max(exp(y) - 1, 0). */
int mrp_rpnn_predict_rows(mrp_ctx* ctx, const float* dev_rows, int n, int precision, float* dev_y);

/* ---------------------------------------------------------------------------------------
 * Volume ingest (VolumeRender::LoadData and its DataCache, core/volume.cu:970-1149).
 * A path containing CLOUD0, CLOUD1, MODEL0 or MODEL1 reads <path>.bin (int32 N, then N^3
 * float32) with that dataset's fixed max density (510.416656, 505.490356, 64, 23.8960018);
 * a path ending in "txt" (N, then N^3 floats) or "vox" (256^3 integers as text, 64 per line,
 * header lines starting with w / h / d, value * 64 / 255) is parsed and written back as
 * <path>.bin (best effort); a path ending in "bin" is read as such.  max_density is then
 * max(grid) (at least 1e-5; the reference leaves it unchanged for txt / vox).
 * --------------------------------------------------------------------------------------- */
/* Host only (no device work): *host_grid is malloc'd, N^3 floats; free with mrp_host_free. */
int mrp_volume_read(const char* path, float** host_grid, int* resolution, float* max_density);
void mrp_host_free(void* p);
/* mrp_volume_read + mrp_volume_create / set_data / update.  The four named datasets' host
 * grids stay cached in the context (the reference's 4-slot DataCache): loading one again
 * reads no file. */
int mrp_volume_load(mrp_ctx* ctx, const char* path, mrp_volume** out, float* max_density);

/* ---------------------------------------------------------------------------------------
 * Environment lighting (SetHDRI / SetEnvExp, core/volume.cu:1698-1736; SkyBox,
 * core/render.cuh:24-27).  With a map set, rays that leave the volume (sky pixels of the
 * predictor, escaping paths of the path tracer) see the map at (atan2(-z, x) / 2pi + 0.5,
 * acos(y) / pi), bilinear, wrapped, times the exposure; without one they see black.
 * --------------------------------------------------------------------------------------- */
/* Radiance RGBE .hdr reader (flat and RLE scanlines, as core/hdr_loader.h): width*height
 * RGBA floats (alpha 0), rows in file order; malloc'd, free with mrp_host_free.  Host only. */
int mrp_hdr_read(const char* path, float** host_rgba, int* width, int* height);
/* host_rgba NULL removes the map. */
int mrp_ctx_set_env_map(mrp_ctx* ctx, const float* host_rgba, int width, int height);
int mrp_ctx_load_hdri(mrp_ctx* ctx, const char* path);
int mrp_ctx_set_env_exp(mrp_ctx* ctx, float exposure); /* default 1 */
/* The reference's per-frame light choice with a map set (core/volume.cu:494-509, 552-568):
 * with probability rate = 4 e / (4 e + max(light_color)) (u < rate, u the caller's uniform
 * draw in [0, 1)) the light becomes direction Roberts2(counter) on the sphere with colour
 * map(direction) * 4 e / rate (clamped bilinear of the host image); otherwise the sun with
 * colour light_color / (1 - rate).  Returns 1 when the environment was chosen, 0 otherwise
 * (and copies the light through when there is no map).  Host only. */
int mrp_env_light(const float* host_rgba, int width, int height, float env_exp, const float light_dir[3],
                  const float light_color[3], int counter, float u, float out_dir[3], float out_color[3]);
int mrp_ctx_env_light(mrp_ctx* ctx, const float light_dir[3], const float light_color[3], int counter, float u,
                      float out_dir[3], float out_color[3]);

/* Delta-tracking path tracer (judge): RenderCamera<false> / CalculateRadiance
 * (core/volume.cu:198-228, core/render.cuh:535-597) with Philox streams keyed
 * (seed, frame) / (pixel, sample) and FP64 per-pixel Welford accumulators.
 * Next-event transmittance (core/render.cuh:589 reads whatever TR mip 0 holds):
 *   nee_baked = 0: the unbiased stochastic Tr (core/render.cuh:60-103) at every vertex;
 *   nee_baked = 1: TR mip 0 of mrp_volume_update_tr (Update_TR, 128^3), the view's light/alpha;
 *   nee_baked = 2: the cached stochastic bake of mrp_volume_bake_cached_tr (Update_CachedTR,
 *     768^3 by default), which the reference's Render(..., PT, ...) uses; it must have been
 *     baked for the view's light_dir, alpha and this tr_scale.
 * dev_mean3 / dev_m2_3: width*height*3 doubles; samples [spp0, spp0+spp) are added to
 * the running (n = spp0 on entry) accumulators. */
/* Update_CachedTR -> Fill_BakeTR (core/volume.cu:1669-1696, 1562-1587): a res^3 grid over the
 * unit box (res = 0: the reference's 768; 2..1024), texel (x, y, z) at ((x, y, z) + 0.5) / res -
 * 0.5 holding the mean of 5 stochastic Tr estimates towards normalize(light_dir) where the
 * density at the texel centre is positive, else of 2.  Each estimate draws from its own Philox
 * stream keyed (seed, 1) / (texel, estimate); the reference re-seeds one cuRAND state inside
 * its loop from the x/y thread and block indices only, so its estimates of a texel repeat and
 * texels differing in z share draws -- same expectation, more variance.  Stored with the
 * volume (res^3 floats, 1.8 GB at 768) until the grid changes or the volume is destroyed;
 * cached on exact (light_dir, alpha, tr_scale, seed, res) like the reference's
 * (lightDir, alpha), so a repeated call is free.  Asynchronous on the context's stream. */
int mrp_volume_bake_cached_tr(mrp_volume* vol, const float light_dir[3], float alpha, float tr_scale, uint32_t seed,
                              int res);
int mrp_volume_download_cached_tr(mrp_volume* vol, float* host); /* res^3 floats; blocking */
/* GetSamples -> GetSampleMulti -> GetSample (core/volume.cu:710-780, 36-47;
 * core/render.cuh:615-654): the radiance arriving at point ori_i (inside the unit box) that
 * dataGen.cu uses as the predictor's training target.  Each path starts in a direction drawn
 * from the HG(g_i) lobe around dir_i, then delta tracking with next-event estimation through the
 * stochastic Tr towards light_i (unit vectors), throughput times 1.001 * scatter_i per vertex,
 * at most multi_scatter vertices, escapes add the environment map.  Path s of task i draws from
 * the Philox stream keyed (seed, 0) / (i, s).  dev_out3: n*3 floats, the mean over spp paths;
 * dev_m2_3 (may be NULL): n*3 doubles, the FP64 sum of squared deviations (Welford), so
 * sqrt(m2 / (spp (spp - 1))) is the mean's standard error.  All arrays are device arrays of n
 * entries (float3 as 3 floats).  Asynchronous. */
int mrp_point_radiance(mrp_ctx* ctx, mrp_volume* vol, int n, const float* dev_ori3, const float* dev_dir3,
                       const float* dev_light3, const float* dev_alpha, const float* dev_g, const float* dev_scatter,
                       const float light_color[3], int multi_scatter, int spp, uint32_t seed, float tr_scale,
                       float* dev_out3, double* dev_m2_3);
/* ShadowTerm_TR / ShadowTerm_TRs (core/render.cuh:656-680, behind VolumeRender::GetTr / GetTrs,
 * core/volume.cu:381-403, 795-830): per task the mean over spp of Tr(ori_i -> light) *
 * HG(dot(dir_i, light), g) with the stochastic Tr towards normalize(light_dir); sample s of task
 * i draws from the Philox stream keyed (seed, 2) / (i, s).  dev_out: n floats.  (GetTrs scales
 * this by lightColor * pi.)  Asynchronous. */
int mrp_tr_samples(mrp_ctx* ctx, mrp_volume* vol, int n, const float* dev_ori3, const float* dev_dir3,
                   const float light_dir[3], float alpha, float g, int spp, uint32_t seed, float tr_scale,
                   float* dev_out);
int mrp_pt_render(mrp_ctx* ctx, mrp_volume* vol, const mrp_view* view, int spp0, int spp, int max_bounce,
                  int nee_baked, float tr_scale, double* dev_mean3, double* dev_m2_3);

/* ---------------------------------------------------------------------------------------
 * Interactive progressive rendering: VolumeRender::Render(float3* target, Histogram*,
 * unsigned* target2, size, ori, up, right, lightDir, lightColor, alpha, multiScatter, g,
 * randseed, rt, toneType, denoise) (core/volume.cu:550-628), called per GUI frame by
 * Camera::Render (core/camera.cu:259-261, core/GUI.hpp:1020-1027).
 *
 * Buffers are caller-owned device memory in the reference layout: dev_target holds 3 floats
 * per pixel (the float3 accumulation image), dev_histo one mrp_histogram per pixel, dev_rgba
 * one 0xAARRGGBB word per pixel; pixel (i, j) is at i*width + j, row i seeing
 * v = (i + jitter) / height (row 0 at the bottom, unlike mrp_render_plain).
 * --------------------------------------------------------------------------------------- */
#define MRP_HISTO_SIZE 10
typedef struct mrp_histogram { /* Histogram (core/volume.hpp:14-19): 132 bytes */
    float bin[3 * MRP_HISTO_SIZE]; /* per channel, 10 bins of the tone-mapped sample c / (c + 1) */
    float total_sample_num;
    float x2; /* running mean of the squared luminance sum */
    float x;  /* running mean of the luminance sum */
} mrp_histogram;
#define MRP_RT_PT 0    /* VolumeRender::RenderType (core/volume.hpp:98-100) */
#define MRP_RT_RPNN 1
#define MRP_RT_MRPNN 2
#define MRP_TONE_NONE 0 /* toneType of Denoise / ReprojectionDenoise (core/volume.cu:332-335) */
#define MRP_TONE_GAMMA 1
#define MRP_TONE_ACES 2
typedef struct mrp_interactive mrp_interactive;
/* The renderer's frame state (VolumeRender's checkboard, last_predict and flip counter, the
 * device exposure, core/volume.hpp:50-51, core/volume.cu:52,481-482) plus a width*height
 * radiance buffer for the frame. */
int mrp_interactive_create(mrp_ctx* ctx, int width, int height, mrp_interactive** out);
void mrp_interactive_destroy(mrp_interactive* it);
int mrp_interactive_set_checkerboard(mrp_interactive* it, int on); /* SetCheckboard; default 1 */
int mrp_interactive_set_exposure(mrp_interactive* it, float exposure); /* SetExposure (volume.cu:1758-1761) */
/* One frame.  render_type MRP_RT_MRPNN: the predictor in `precision` (MRP_FP32 = the
 * reference's arithmetic), max(0, rgb) into the target (on half the pixels in checkerboard
 * mode), then ReprojectionDenoise: the histogram's (total, x, x2) take the target's (r, g, b)
 * and the target is tone-mapped.  MRP_RT_PT: one delta-tracking sample per pixel (baked
 * TR NEE, max_bounce = multiScatter), blended into the target with weight 1 / (1 + n),
 * n = randseed (randseed / 2 with checkerboard), and counted into the histogram; then
 * Denoise: at randseed 0 the checkerboard's other half is filled with the mean of its four
 * neighbours, later frames with denoise on replace pixels whose luminance variance exceeds
 * 0.01 by an 11x11 average weighted by histogram similarity.  Tone mapping: exposure, then
 * none / Gamma 1/2.2 / ACES (core/render.cuh:1144-1236), saturate, 8 bits per channel.
 * The view's light, alpha and g update the LUT and TR mips as the reference does; the
 * jitter stream is keyed by (view->seed, randseed).  Asynchronous, except MRP_RT_RPNN:
 * the RPNN network (MRP_FP32 or MRP_BF16; other precisions are refused with MRP_ERR_INVALID,
 * as by mrp_render_rpnn) in place of the predictor, which synchronises with its march as
 * mrp_render_rpnn. */
int mrp_render_interactive(mrp_interactive* it, mrp_volume* vol, const mrp_view* view, int render_type, int precision,
                           int max_bounce, int randseed, int tone_type, int denoise, float* dev_target,
                           mrp_histogram* dev_histo, uint32_t* dev_rgba);
/* Copies the last frame's radiance, max(0, rgb) and the distance (0 for PT), per pixel in the
 * target layout into dev_out4 (width*height*4 floats); w = -2 for pixels the frame did not
 * render (checkerboard).  No counterpart in the reference.  Asynchronous. */
int mrp_interactive_radiance(mrp_interactive* it, float* dev_out4);

/* ---------------------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, which renders one blocking view per
 * Render call, core/volume.cu:487-548).  Volumes are registered with the server and made
 * device-resident on demand when their grid + pyramids + padded copies (mrp_volume_bytes)
 * fit in free HBM; idle resident volumes are evicted least recently used first to make
 * room.  Requests are batched as they arrive: while up to max_inflight batches run on the
 * GPU, the next one is formed from the requests that arrived meanwhile.  Each request's
 * pixels are computed exactly as when it is rendered alone.
 * --------------------------------------------------------------------------------------- */
typedef struct mrp_server mrp_server;
typedef struct mrp_server_config {
    int32_t precision;        /* MRP_BF16 / MRP_FP8 / MRP_BF16_1T / MRP_FP32 (batched path) */
    int32_t plain;            /* 1: one plain per-frame FP32 launch per request (the baseline) */
    int32_t max_batch_views;  /* requests per batch */
    int64_t max_batch_pixels; /* a batch closes once it holds this many pixels */
    int32_t max_inflight;     /* batches on the GPU at once */
    float g;                  /* Henyey-Greenstein g of every request (one LUT) */
    int64_t hbm_limit_bytes;  /* > 0: admit against this budget instead of hipMemGetInfo */
    int64_t hbm_reserve_bytes;/* kept free beside every admitted volume */
} mrp_server_config;
typedef struct mrp_request {
    int32_t volume_id;        /* from mrp_server_add_volume / mrp_server_add_fbm_volume */
    mrp_view view;            /* volume_index is ignored */
    float* dev_out4;          /* width*height float4, caller-owned, valid until completion */
    uint64_t user_tag;
} mrp_request;
typedef struct mrp_completion {
    uint64_t user_tag;
    double t_submit, t_launch, t_done; /* host steady-clock seconds */
    int32_t status;
} mrp_completion;
typedef struct mrp_server_stats {
    int64_t submitted, completed, pixels, batches, batched_views, admissions, evictions;
    int64_t resident_volumes, resident_bytes, pending;
} mrp_server_stats;
typedef struct mrp_load_spec {
    double rate_rps;          /* Poisson arrival rate */
    double duration_s;        /* arrivals over [0, duration_s) */
    uint64_t seed;
    int32_t n_sizes;          /* 1..8 resolutions, drawn uniformly */
    int32_t sizes[16];        /* (width, height) pairs */
    int32_t buffers_per_size; /* output buffers per resolution (default 16) */
} mrp_load_spec;
typedef struct mrp_load_report {
    int64_t requests, completed, pixels;
    double wall_s, offered_s, mpix_s;
    double p50_ms, p99_ms, max_ms; /* arrival to completion */
} mrp_load_report;

int mrp_server_create(mrp_ctx* ctx, const mrp_server_config* cfg, mrp_server** out);
int mrp_server_destroy(mrp_server* srv);
/* Registers a volume (the server keeps a host copy for re-admission after eviction). */
int mrp_server_add_volume(mrp_server* srv, int N, const float* host_grid, float max_density, int* volume_id);
/* Registers a seeded fBm volume (regenerated on the device at admission). */
int mrp_server_add_fbm_volume(mrp_server* srv, int N, int octaves, float base_freq, uint32_t seed, int* volume_id);
/* Queues a request (thread-safe with respect to mrp_server_step). */
int mrp_server_submit(mrp_server* srv, const mrp_request* req);
/* One scheduling round: reap finished batches, admit volumes, launch batches. Non-blocking. */
int mrp_server_step(mrp_server* srv);
/* Steps until every submitted request has completed. */
int mrp_server_drain(mrp_server* srv);
/* Moves up to max finished requests into out. */
int mrp_server_poll(mrp_server* srv, mrp_completion* out, int max, int* n);
int mrp_server_get_stats(mrp_server* srv, mrp_server_stats* out);
int mrp_server_is_resident(mrp_server* srv, int volume_id);
/* Open-loop load over the registered volumes: Poisson arrivals, random orbit (radius 2,
 * 40 degree field of view), sun on the upper hemisphere, alpha in [16, 128] and resolution
 * from spec->sizes; latency is measured from arrival to completion.  Blocking. */
int mrp_server_run_poisson(mrp_server* srv, const mrp_load_spec* spec, mrp_load_report* report);

#ifdef __cplusplus
}
#endif
#endif /* MRP_C_API_H */
