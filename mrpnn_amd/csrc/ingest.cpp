// Volume ingest and environment lighting:
//   VolumeRender::LoadData with its DataCache (core/volume.cu:970-1149): the four named
//   datasets read from <path>.bin with fixed max densities, .txt and .vox grids parsed and
//   written back as <path>.bin;
//   SetHDRI / SetEnvExp (core/volume.cu:1698-1736) over the Radiance RGBE reader of
//   core/hdr_loader.h;
//   the per-frame choice between the sun and one direction of the environment map
//   (core/volume.cu:494-509, 552-568): Roberts2 (core/vector.cu:179-183), UniformSampleSphere
//   and Image_host::Sample (core/volume.hpp:23-43).
// Host code only; the device side of the environment map is sky_box() in dev_common.hpp.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/mrp_c_api.h"
#include "capi_internal.h"

namespace {

struct NamedSet {
    const char* tag;
    float max_density;
};
// core/volume.cu:979-1003
const NamedSet kNamed[4] = {{"CLOUD0", 510.416656f}, {"CLOUD1", 505.490356f}, {"MODEL0", 64.0000000f},
                            {"MODEL1", 23.8960018f}};

int named_slot(const char* path) {
    for (int i = 0; i < 4; ++i)
        if (strstr(path, kNamed[i].tag)) return i;
    return -1;
}

bool read_file(const std::string& path, std::vector<char>& buf) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(n > 0 ? (size_t)n + 1 : 1);
    size_t got = n > 0 ? fread(buf.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    buf[got] = 0;
    buf.resize(got + 1);
    return n >= 0 && got == (size_t)n;
}

// .bin: int32 N, then N^3 float32 (core/volume.cu:1011-1035)
int parse_bin(const std::string& path, std::vector<float>& grid, int& N) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return mrp_fail(MRP_ERR_IO, "cannot open %s", path.c_str());
    int32_t n = 0;
    if (fread(&n, 4, 1, f) != 1 || n < 2 || n > 2048) {
        fclose(f);
        return mrp_fail(MRP_ERR_IO, "bad volume header in %s", path.c_str());
    }
    // the header may not size an allocation the file cannot fill
    const size_t want = (size_t)n * n * n;
    long size = -1;
    if (fseek(f, 0, SEEK_END) == 0) size = ftell(f);
    if (size < 0 || (size_t)size < 4 + want * sizeof(float) || fseek(f, 4, SEEK_SET) != 0) {
        fclose(f);
        return mrp_fail(MRP_ERR_IO, "truncated volume file %s (%d^3 floats declared)", path.c_str(), (int)n);
    }
    grid.resize(want);
    size_t got = fread(grid.data(), sizeof(float), grid.size(), f);
    fclose(f);
    if (got != grid.size()) return mrp_fail(MRP_ERR_IO, "truncated volume file %s", path.c_str());
    N = n;
    return MRP_OK;
}

// .txt: the resolution, then N^3 whitespace-separated floats (core/volume.cu:1067-1084)
int parse_txt(const std::string& path, std::vector<float>& grid, int& N) {
    std::vector<char> buf;
    if (!read_file(path, buf)) return mrp_fail(MRP_ERR_IO, "cannot read %s", path.c_str());
    char* p = buf.data();
    char* e = nullptr;
    long n = strtol(p, &e, 10);
    if (e == p || n < 2 || n > 2048) return mrp_fail(MRP_ERR_IO, "bad resolution in %s", path.c_str());
    // every value takes at least one character and one separator
    if ((size_t)n * n * n > buf.size() / 2 + 1)
        return mrp_fail(MRP_ERR_IO, "%s: %ld^3 values declared in %zu bytes", path.c_str(), n, buf.size() - 1);
    p = e;
    grid.resize((size_t)n * n * n);
    for (size_t i = 0; i < grid.size(); ++i) {
        float v = strtof(p, &e);
        if (e == p) return mrp_fail(MRP_ERR_IO, "%s: %zu of %zu values", path.c_str(), i, grid.size());
        grid[i] = v;
        p = e;
    }
    N = (int)n;
    return MRP_OK;
}

// .vox: 256^3 bytes as text, 64 integers per line, lines starting with w / h / d are the
// header; value * 64 / 255 (core/volume.cu:1085-1110)
int parse_vox(const std::string& path, std::vector<float>& grid, int& N) {
    std::vector<char> buf;
    if (!read_file(path, buf)) return mrp_fail(MRP_ERR_IO, "cannot read %s", path.c_str());
    const size_t total = (size_t)256 * 256 * 256;
    grid.assign(total, 0.0f);
    const float inv = 64.0 / 255.0;
    size_t i = 0;
    char* p = buf.data();
    char* end = buf.data() + buf.size() - 1;
    while (p < end && i < total) {
        char* eol = (char*)memchr(p, '\n', (size_t)(end - p));
        if (!eol) eol = end;
        const char first = *p;
        if (first != 'w' && first != 'h' && first != 'd' && eol > p && first != '\r') {
            char* q = p;
            for (int j = 0; j < 64 && i + j < total; ++j) {
                char* e = nullptr;
                long d = strtol(q, &e, 10);
                if (e == q || e > eol) return mrp_fail(MRP_ERR_IO, "%s: short line near value %zu", path.c_str(), i);
                grid[i + j] = (float)d * inv;
                q = e;
            }
            i += 64;
        }
        p = eol + 1;
    }
    if (i < total) return mrp_fail(MRP_ERR_IO, "%s: %zu of %zu values", path.c_str(), i, total);
    N = 256;
    return MRP_OK;
}

void write_bin(const std::string& path, const std::vector<float>& grid, int N) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return;  // a cache: failing to write it leaves the load itself intact
    int32_t n = N;
    fwrite(&n, 4, 1, f);
    fwrite(grid.data(), sizeof(float), grid.size(), f);
    fclose(f);
}

// Reads the grid of `path` as LoadData does; max_density as LoadData sets it for the named
// datasets, max(grid) (at least 1e-5) otherwise.
int read_volume(const char* path, bool write_cache, std::vector<float>& grid, int& N, float& md) {
    const std::string p(path);
    const int slot = named_slot(path);
    int rc;
    if (slot >= 0) {
        rc = parse_bin(p + ".bin", grid, N);
        if (rc) return rc;
        md = kNamed[slot].max_density;
        return MRP_OK;
    }
    const std::string ext = p.size() >= 3 ? p.substr(p.size() - 3) : std::string();
    if (ext == "txt") rc = parse_txt(p, grid, N);
    else if (ext == "vox") rc = parse_vox(p, grid, N);
    else if (ext == "bin") rc = parse_bin(p, grid, N);
    else return mrp_fail(MRP_ERR_IO, "File not found: %s (not a named dataset, .txt, .vox or .bin)", path);
    if (rc) return rc;
    md = 0.00001f;
    for (float v : grid) md = v > md ? v : md;
    if (write_cache && (ext == "txt" || ext == "vox")) write_bin(p + ".bin", grid, N);  // volume.cu:1143-1148
    return MRP_OK;
}

// Radiance RGBE ".hdr" (the format core/hdr_loader.h reads): "#?" magic, header lines,
// "-Y <h> +X <w>", then scanlines, each either flat RGBE or new-style RLE (2, 2, len).
bool rgbe_scanline(const unsigned char*& p, const unsigned char* end, unsigned char* line, unsigned w) {
    auto flat = [&](unsigned char* dst, unsigned n) {
        if ((size_t)(end - p) < (size_t)n * 4) return false;
        memcpy(dst, p, (size_t)n * 4);
        p += (size_t)n * 4;
        return true;
    };
    const bool rle_ok = w >= 8 && w <= 0x7fff;
    if (!rle_ok || end - p < 4 || p[0] != 2) return flat(line, w);
    if (p[1] != 2 || (p[2] & 128)) return flat(line, w);  // old-style pixel starting with 2
    const unsigned len = (unsigned)p[2] << 8 | p[3];
    if (len != w) return false;
    p += 4;
    for (int c = 0; c < 4; ++c) {
        for (unsigned pos = 0; pos < w;) {
            if (p >= end) return false;
            unsigned num = *p++;
            if (num > 128) {
                num &= 127;
                if (p >= end || pos + num > w) return false;
                const unsigned char v = *p++;
                for (unsigned j = 0; j < num; ++j) line[(pos++) * 4 + c] = v;
            } else {
                if ((size_t)(end - p) < num || pos + num > w) return false;
                for (unsigned j = 0; j < num; ++j) line[(pos++) * 4 + c] = *p++;
            }
        }
    }
    return true;
}

float rgbe_scale(unsigned char e) {  // 2^(e - 136) through the exponent bits, as the reader
    union {
        float f;
        uint32_t i;
    } u;
    u.i = (uint32_t)(((int)e - 9) << 23) & 0x7f800000u;
    return u.f;
}

int read_hdr(const char* path, std::vector<float>& rgba, int& w, int& h) {
    std::vector<char> buf;
    if (!read_file(path, buf)) return mrp_fail(MRP_ERR_IO, "cannot read %s", path);
    const char* s = buf.data();
    const char* end = s + buf.size() - 1;
    if (end - s < 2 || s[0] != '#' || s[1] != '?') return mrp_fail(MRP_ERR_IO, "%s: not a Radiance file", path);
    unsigned rx = 0, ry = 0;
    const char* line = s;
    for (;;) {
        const char* eol = (const char*)memchr(line, '\n', (size_t)(end - line));
        if (!eol) return mrp_fail(MRP_ERR_IO, "%s: header without resolution", path);
        std::string l(line, eol);
        line = eol + 1;
        if (l.rfind("FORMAT=", 0) == 0 && l.compare(7, 15, "32-bit_rle_rgbe") != 0 &&
            l.compare(7, 15, "32-bit_rle_xyze") != 0)
            return mrp_fail(MRP_ERR_IO, "%s: unknown format %s", path, l.c_str() + 7);
        if (!l.empty() && (l[0] == '-' || l[0] == '+')) {
            const size_t x = l.find('X'), y = l.find('Y');
            if (x == std::string::npos || y == std::string::npos || sscanf(l.c_str() + x + 1, "%u", &rx) != 1 ||
                sscanf(l.c_str() + y + 1, "%u", &ry) != 1 || rx == 0 || ry == 0 || rx > 65536 || ry > 65536)
                return mrp_fail(MRP_ERR_IO, "%s: bad resolution line", path);
            break;
        }
    }
    if ((uint64_t)rx * ry > ((uint64_t)1 << 26))  // the limit of mrp_ctx_set_env_map
        return mrp_fail(MRP_ERR_IO, "%s: %u x %u pixels exceed 2^26", path, rx, ry);
    const unsigned char* p = (const unsigned char*)line;
    const unsigned char* pend = (const unsigned char*)end;
    rgba.assign((size_t)rx * ry * 4, 0.0f);
    std::vector<unsigned char> sl((size_t)rx * 4);
    for (unsigned j = 0; j < ry; ++j) {
        if (!rgbe_scanline(p, pend, sl.data(), rx)) return mrp_fail(MRP_ERR_IO, "%s: bad scanline %u", path, j);
        for (unsigned i = 0; i < rx; ++i) {
            const unsigned char* q = &sl[(size_t)i * 4];
            float* o = &rgba[((size_t)j * rx + i) * 4];
            if (q[3] == 0) continue;
            const float f = rgbe_scale(q[3]);
            o[0] = (float)(q[0] + 0.5f) * f;
            o[1] = (float)(q[1] + 0.5f) * f;
            o[2] = (float)(q[2] + 0.5f) * f;
        }
    }
    w = (int)rx;
    h = (int)ry;
    return MRP_OK;
}

struct F3 {
    float x, y, z;
};
F3 f3lerp(F3 a, F3 b, float v) {
    return {a.x * (1.0f - v) + b.x * v, a.y * (1.0f - v) + b.y * v, a.z * (1.0f - v) + b.z * v};
}
// Image_host::Sample: bilinear with clamped addressing (core/volume.hpp:23-43)
F3 image_sample(const float* rgba, int sx, int sy, float u, float v) {
    auto px = [&](int x, int y) {
        x = x < 0 ? 0 : (x > sx - 1 ? sx - 1 : x);
        y = y < 0 ? 0 : (y > sy - 1 ? sy - 1 : y);
        const float* q = rgba + ((size_t)y * sx + x) * 4;
        return F3{q[0], q[1], q[2]};
    };
    const float fx = u * (float)sx - 0.5f, fy = v * (float)sy - 0.5f;
    const int x = (int)floorf(fx), y = (int)floorf(fy);
    const float wx = fx - (float)x, wy = fy - (float)y;
    return f3lerp(f3lerp(px(x, y), px(x + 1, y), wx), f3lerp(px(x, y + 1), px(x + 1, y + 1), wx), wy);
}

// Host allocations of the readers may fail (std::bad_alloc); nothing may unwind through the C ABI.
template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return mrp_fail(MRP_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return mrp_fail(MRP_ERR_IO, "%s", e.what());
    }
}

}  // namespace

extern "C" int mrp_volume_read(const char* path, float** host_grid, int* resolution, float* max_density) {
    return guarded([&]() -> int {
        if (!path || !host_grid || !resolution || !max_density) return mrp_fail(MRP_ERR_INVALID, "null argument");
        std::vector<float> grid;
        int N = 0;
        float md = 0.0f;
        int rc = read_volume(path, true, grid, N, md);
        if (rc) return rc;
        float* out = (float*)malloc(grid.size() * sizeof(float));
        if (!out) return mrp_fail(MRP_ERR_NOMEM, "host allocation failed");
        memcpy(out, grid.data(), grid.size() * sizeof(float));
        *host_grid = out;
        *resolution = N;
        *max_density = md;
        return MRP_OK;
    });
}

extern "C" void mrp_host_free(void* p) { free(p); }

extern "C" int mrp_volume_load(mrp_ctx* c, const char* path, mrp_volume** out, float* max_density) {
    return guarded([&]() -> int {
        if (!c || !path || !out) return mrp_fail(MRP_ERR_INVALID, "null argument");
        std::vector<float> local;
        const std::vector<float>* grid = &local;
        int N = 0;
        float md = 0.0f;
        const int slot = named_slot(path);
        if (slot >= 0 && c->data_cache_n[slot] > 0) {  // DataCache hit (core/volume.cu:1008-1053)
            grid = &c->data_cache[slot];
            N = c->data_cache_n[slot];
            md = kNamed[slot].max_density;
        } else {
            int rc = read_volume(path, true, local, N, md);
            if (rc) return rc;
            if (slot >= 0) {
                c->data_cache[slot].swap(local);
                c->data_cache_n[slot] = N;
                grid = &c->data_cache[slot];
            }
        }
        mrp_volume* v = nullptr;
        int rc = mrp_volume_create(c, N, &v);
        if (!rc) rc = mrp_volume_set_data(v, grid->data(), md);
        if (!rc) rc = mrp_volume_update(v);
        if (rc) {
            if (v) mrp_volume_destroy(v);
            return rc;
        }
        *out = v;
        if (max_density) *max_density = md;
        return MRP_OK;
    });
}

extern "C" int mrp_hdr_read(const char* path, float** host_rgba, int* width, int* height) {
    return guarded([&]() -> int {
        if (!path || !host_rgba || !width || !height) return mrp_fail(MRP_ERR_INVALID, "null argument");
        std::vector<float> px;
        int w = 0, h = 0;
        int rc = read_hdr(path, px, w, h);
        if (rc) return rc;
        float* o = (float*)malloc(px.size() * sizeof(float));
        if (!o) return mrp_fail(MRP_ERR_NOMEM, "host allocation failed");
        memcpy(o, px.data(), px.size() * sizeof(float));
        *host_rgba = o;
        *width = w;
        *height = h;
        return MRP_OK;
    });
}

extern "C" int mrp_ctx_set_env_map(mrp_ctx* c, const float* host_rgba, int width, int height) {
    return guarded([&]() -> int {
        if (!c) return mrp_fail(MRP_ERR_INVALID, "null argument");
        if (c->d_env) {
            if (hipStreamSynchronize(c->stream) != hipSuccess) return mrp_fail(MRP_ERR_HIP, "stream sync failed");
            (void)hipFree(c->d_env);
            c->d_env = nullptr;
        }
        c->env_host.clear();
        c->env_w = c->env_h = 0;
        if (!host_rgba) return MRP_OK;  // no environment: sky radiance 0
        if (width < 1 || height < 1 || (int64_t)width * height > (1 << 26))
            return mrp_fail(MRP_ERR_INVALID, "environment map %d x %d out of range", width, height);
        const size_t n = (size_t)width * height * 4;
        if (hipMalloc(&c->d_env, n * sizeof(float)) != hipSuccess) {
            c->d_env = nullptr;
            return mrp_fail(MRP_ERR_NOMEM, "environment map allocation failed");
        }
        if (hipMemcpy(c->d_env, host_rgba, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(c->d_env);
            c->d_env = nullptr;
            return mrp_fail(MRP_ERR_HIP, "environment map upload failed");
        }
        c->env_host.assign(host_rgba, host_rgba + n);
        c->env_w = width;
        c->env_h = height;
        return MRP_OK;
    });
}

extern "C" int mrp_ctx_load_hdri(mrp_ctx* c, const char* path) {
    return guarded([&]() -> int {
        if (!c || !path) return mrp_fail(MRP_ERR_INVALID, "null argument");
        std::vector<float> px;
        int w = 0, h = 0;
        int rc = read_hdr(path, px, w, h);
        if (rc) return rc;
        return mrp_ctx_set_env_map(c, px.data(), w, h);
    });
}

extern "C" int mrp_ctx_set_env_exp(mrp_ctx* c, float exposure) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "null argument");
    c->env_exp = exposure;
    return MRP_OK;
}

extern "C" int mrp_env_light(const float* host_rgba, int width, int height, float env_exp, const float light_dir[3],
                             const float light_color[3], int counter, float u, float out_dir[3], float out_color[3]) {
    if (!light_dir || !light_color || !out_dir || !out_color) return mrp_fail(MRP_ERR_INVALID, "null argument");
    for (int k = 0; k < 3; ++k) {
        out_dir[k] = light_dir[k];
        out_color[k] = light_color[k];
    }
    if (!host_rgba || width < 1 || height < 1) return 0;
    const float lmax = std::max(light_color[0], std::max(light_color[1], light_color[2]));
    const float rate = env_exp * 4 / (env_exp * 4 + lmax);
    if (u < rate) {
        // Roberts2(n) = frac((1/g, 1/g^2) n + 0.5), g the plastic number (core/vector.cu:179-183)
        const float g = 1.32471795724474602596;
        const float ax = 1.0f / g, ay = 1.0f / (g * g);
        const float rx = ax * (float)counter + 0.5f, ry = ay * (float)counter + 0.5f;
        const float ex = rx - (float)(long)rx, ey = ry - (float)(long)ry;  // frac (core/vector.cu:20)
        // UniformSampleSphere (core/vector.cu:195-208)
        const float phi = 2.0f * 3.14159265359f * ex;
        const float ct = 1.0f - 2.0f * ey;
        const float st = sqrtf(1.0f - ct * ct);
        const float d[3] = {st * cosf(phi), st * sinf(phi), ct};
        const float uu = atan2f(-d[2], d[0]) * (float)(0.5 / 3.1415926) + 0.5f;
        const float vv = acosf(fmaxf(fminf(d[1], 1.0f), -1.0f)) * (float)(1.0 / 3.1415926);
        const F3 c = image_sample(host_rgba, width, height, uu, vv);
        const float s = env_exp * 4;
        for (int k = 0; k < 3; ++k) out_dir[k] = d[k];
        out_color[0] = c.x * s / rate;
        out_color[1] = c.y * s / rate;
        out_color[2] = c.z * s / rate;
        return 1;
    }
    for (int k = 0; k < 3; ++k) out_color[k] = light_color[k] / (1 - rate);
    return 0;
}

extern "C" int mrp_ctx_env_light(mrp_ctx* c, const float light_dir[3], const float light_color[3], int counter, float u,
                                 float out_dir[3], float out_color[3]) {
    if (!c) return mrp_fail(MRP_ERR_INVALID, "null argument");
    return mrp_env_light(c->env_host.empty() ? nullptr : c->env_host.data(), c->env_w, c->env_h, c->env_exp, light_dir,
                         light_color, counter, u, out_dir, out_color);
}
