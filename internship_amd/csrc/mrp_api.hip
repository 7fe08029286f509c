// mrp_api.hip - C ABI: errors, device-memory plumbing, volume lifetime.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>
#include <string>

#include "mrp_frame.h"
#include "mrp_internal.h"

namespace mrp {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    return fail(MRP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

EnvMap env_map(const Volume* v, vec3 sky) {
    EnvMap e;
    e.px = v->hdri;
    e.w = v->hdri_w;
    e.h = v->hdri_h;
    e.exposure = v->hdri_exp;
    e.sky = sky;
    return e;
}

void env_light_of_sample(const Volume* v, int n, uint64_t seed, mrp_vec3* light_dir, mrp_vec3* light_color) {
    if (!v->hdri) return;
    EnvMap e = env_map(v, v3(0, 0, 0));
    e.px = v->hdri_host.data();  // the host copy
    Rng rng = rng_make(0u, (uint32_t)n, 6u, seed);
    const float choose = rng_next(&rng);
    vec3 d = v3(light_dir->x, light_dir->y, light_dir->z), c = v3(light_color->x, light_color->y, light_color->z);
    env_light(e, n, choose, &d, &c);
    *light_dir = mrp_vec3{d.x, d.y, d.z};
    *light_color = mrp_vec3{c.x, c.y, c.z};
}

int ensure_samples(Volume* v, size_t bytes) {
    if (v->samples_bytes >= bytes) return MRP_OK;
    if (v->samples) {
        MRP_HIP(hipStreamSynchronize(v->stream));
        MRP_HIP(hipFree(v->samples));
        v->samples = nullptr;
        v->samples_bytes = 0;
    }
    hipError_t e = hipMalloc(&v->samples, bytes);
    if (e != hipSuccess) {
        v->samples = nullptr;
        return fail(MRP_E_NOMEM, std::string("sample buffer: ") + hipGetErrorString(e));
    }
    v->samples_bytes = bytes;
    return MRP_OK;
}

}  // namespace mrp

using namespace mrp;

extern "C" {

const char* mrp_last_error(void) { return g_last_error.c_str(); }

int mrp_device_count(int* count) {
    if (!count) return fail(MRP_E_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return MRP_OK;
}

int mrp_malloc(void** dev, size_t bytes) {
    if (!dev) return fail(MRP_E_INVALID, "dev is NULL");
    *dev = nullptr;
    if (bytes == 0) return MRP_OK;
    hipError_t e = hipMalloc(dev, bytes);
    if (e != hipSuccess) {
        *dev = nullptr;
        return fail(MRP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return MRP_OK;
}
int mrp_free(void* dev) {
    if (dev) MRP_HIP(hipFree(dev));
    return MRP_OK;
}
int mrp_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return MRP_OK;
    if (!dst || !src) return fail(MRP_E_INVALID, "NULL pointer");
    MRP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MRP_OK;
}
int mrp_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return MRP_OK;
    if (!dst || !src) return fail(MRP_E_INVALID, "NULL pointer");
    MRP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MRP_OK;
}
int mrp_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return MRP_OK;
    if (!dst || !src) return fail(MRP_E_INVALID, "NULL pointer");
    MRP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MRP_OK;
}
// dst row r = src row idx[r] (16-byte words); an index past n_src gives a zero row.
__global__ void gather_rows_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                   const uint32_t* __restrict__ idx, uint32_t n_rows, uint32_t words,
                                   uint32_t n_src) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_rows * words) return;
    const uint32_t r = (uint32_t)(t / words), w = (uint32_t)(t - (size_t)r * words);
    const uint32_t s = idx[r];
    dst[t] = s < n_src ? src[(size_t)s * words + w] : make_uint4(0u, 0u, 0u, 0u);
}

int mrp_gather_rows(void* dst_dev, const void* src_dev, const uint32_t* idx_dev, int n_rows, int row_bytes,
                    int n_src_rows) {
    if (n_rows < 0 || n_src_rows < 0) return fail(MRP_E_INVALID, "negative row count");
    if (n_rows == 0) return MRP_OK;
    if (!dst_dev || !src_dev || !idx_dev) return fail(MRP_E_INVALID, "NULL pointer");
    if (row_bytes <= 0 || row_bytes % 16) return fail(MRP_E_INVALID, "row_bytes must be a positive multiple of 16");
    if (((uintptr_t)dst_dev | (uintptr_t)src_dev) % 16) return fail(MRP_E_INVALID, "rows must be 16-byte aligned");
    const uint32_t words = (uint32_t)(row_bytes / 16);
    const size_t total = (size_t)n_rows * words;
    gather_rows_kernel<<<(unsigned)((total + 255) / 256), 256>>>((uint4*)dst_dev, (const uint4*)src_dev, idx_dev,
                                                                 (uint32_t)n_rows, words, (uint32_t)n_src_rows);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

int mrp_memset_zero(void* dev, size_t bytes) {
    if (bytes == 0) return MRP_OK;
    if (!dev) return fail(MRP_E_INVALID, "NULL pointer");
    MRP_HIP(hipMemset(dev, 0, bytes));
    return MRP_OK;
}

int mrp_tone_map_host(const float* rgb, size_t n, int tone, float* out) {
    if ((!rgb || !out) && n) return fail(MRP_E_INVALID, "NULL pointer");
    if (tone < 0 || tone > 2) return fail(MRP_E_INVALID, "tone must be 0 (none), 1 (gamma) or 2 (ACES)");
    for (size_t i = 0; i < n; ++i) {
        vec3 c = v3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
        if (tone == 1) c = tone_gamma(c);
        else if (tone == 2) c = tone_aces(c);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
    return MRP_OK;
}

int mrp_volume_create(int resolution, mrp_volume** out) {
    if (!out) return fail(MRP_E_INVALID, "out is NULL");
    *out = nullptr;
    if (resolution < 2 || resolution > 1024) return fail(MRP_E_INVALID, "resolution must be in [2, 1024]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(MRP_E_HIP, "no HIP device");
    mrp_volume* h = new (std::nothrow) mrp_volume();
    if (!h) return fail(MRP_E_NOMEM, "host allocation failed");
    Volume* v = &h->v;
    v->res = resolution;
    int rc = MRP_OK;
    auto bail = [&](int code, const std::string& m) {
        rc = fail(code, m);
        mrp_volume_destroy(h);
        return rc;
    };
    if (hipGetDevice(&v->device) != hipSuccess) return bail(MRP_E_HIP, "hipGetDevice");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, v->device) != hipSuccess) return bail(MRP_E_HIP, "hipGetDeviceProperties");
    v->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(MRP_E_HIP, "hipStreamCreate");
    if (hipMalloc(&v->grid, padded_count(resolution) * sizeof(float)) != hipSuccess)
        return bail(MRP_E_NOMEM, "grid allocation failed");
    if (hipMalloc(&v->cells, cell_count(resolution) * 2 * sizeof(float4)) != hipSuccess)
        return bail(MRP_E_NOMEM, "cell grid allocation failed");
    if (hipMalloc(&v->quads, quad_count(resolution) * sizeof(float4)) != hipSuccess)
        return bail(MRP_E_NOMEM, "quad grid allocation failed");
    if (hipMalloc(&v->occ, occupancy_words(resolution) * sizeof(uint32_t)) != hipSuccess)
        return bail(MRP_E_NOMEM, "occupancy allocation failed");
    for (int m = 0; m < kNumMips; ++m)
        if (hipMalloc(&v->mips[m], padded_count(mip_res(m)) * sizeof(float)) != hipSuccess)
            return bail(MRP_E_NOMEM, "mip allocation failed");
    for (int m = 0; m < kNumTrMips; ++m) {
        const size_t r = (size_t)tr_mip_res(m);
        if (hipMalloc(&v->tr_mips[m], r * r * r * sizeof(float)) != hipSuccess)
            return bail(MRP_E_NOMEM, "TR mip allocation failed");
    }
    if (hipMalloc(&v->hg_lut, (size_t)kLutSize * kLutSize * sizeof(float)) != hipSuccess)
        return bail(MRP_E_NOMEM, "HG LUT allocation failed");
    for (int m = 0; m < kNumMips; ++m)
        if (hipMalloc(&v->mip_cells[m], cell_count(mip_res(m)) * 2 * sizeof(float4)) != hipSuccess)
            return bail(MRP_E_NOMEM, "mip cell allocation failed");
    for (int m = 0; m < kNumTrMips; ++m)
        if (hipMalloc(&v->tr_cells[m], clamp_cell_count(tr_mip_res(m)) * 2 * sizeof(float4)) != hipSuccess)
            return bail(MRP_E_NOMEM, "TR cell allocation failed");
    if (hipMalloc(&v->lut_cells, lut_cell_count(kLutSize, kLutSize) * sizeof(float4)) != hipSuccess)
        return bail(MRP_E_NOMEM, "HG LUT cell allocation failed");
    if (hipHostMalloc((void**)&v->host_words, 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return bail(MRP_E_NOMEM, "pinned host words");
    if (hipMalloc(&v->mip_cells16, (size_t)mip16_offset(kNumMips) * sizeof(uint4)) != hipSuccess)
        return bail(MRP_E_NOMEM, "FP16 mip cell allocation failed");
    if (hipMalloc(&v->tr_cells16, (size_t)tr16_offset(kNumTrMips) * sizeof(uint4)) != hipSuccess)
        return bail(MRP_E_NOMEM, "FP16 TR cell allocation failed");
    if (hipMalloc(&v->queues, kNumQueues * sizeof(uint32_t)) != hipSuccess) return bail(MRP_E_NOMEM, "queues");
    if (hipMalloc(&v->counters, 4 * sizeof(unsigned long long)) != hipSuccess) return bail(MRP_E_NOMEM, "counters");
    *out = h;
    return MRP_OK;
}

int mrp_volume_set_hdri(mrp_volume* vol, const float* rgba_host, int width, int height, float exposure) {
    if (!vol) return fail(MRP_E_INVALID, "NULL volume");
    Volume* v = &vol->v;
    MRP_HIP(hipSetDevice(v->device));
    if (width <= 0 || height <= 0 || !rgba_host) {  // back to the constant sky
        if (v->hdri) {
            MRP_HIP(hipStreamSynchronize(v->stream));
            MRP_HIP(hipFree(v->hdri));
        }
        v->hdri = nullptr;
        v->hdri_w = v->hdri_h = 0;
        v->hdri_host.clear();
        return MRP_OK;
    }
    if ((long long)width * height > (1ll << 28)) return fail(MRP_E_INVALID, "environment map too large");
    if (!(exposure >= 0.0f)) return fail(MRP_E_INVALID, "exposure must be >= 0");
    const size_t n = 4 * (size_t)width * height;
    MRP_HIP(hipStreamSynchronize(v->stream));
    if (v->hdri && (size_t)v->hdri_w * v->hdri_h != (size_t)width * height) {
        MRP_HIP(hipFree(v->hdri));
        v->hdri = nullptr;
    }
    if (!v->hdri && hipMalloc(&v->hdri, n * sizeof(float)) != hipSuccess) {
        v->hdri = nullptr;
        v->hdri_w = v->hdri_h = 0;
        v->hdri_host.clear();
        return fail(MRP_E_NOMEM, "environment map allocation failed");
    }
    MRP_HIP(hipMemcpy(v->hdri, rgba_host, n * sizeof(float), hipMemcpyHostToDevice));
    v->hdri_host.assign(rgba_host, rgba_host + n);
    v->hdri_w = width;
    v->hdri_h = height;
    v->hdri_exp = exposure;
    return MRP_OK;
}

// Radiance RGBE (".hdr") reader: the header lines up to a blank line (FORMAT
// must be 32-bit_rle_rgbe if given), the resolution line "-Y h +X w", then h
// scanlines, each flat (4 bytes per pixel) or new-style run-length encoded
// (0x02 0x02 then the width, then the 4 channels one after another in runs:
// a count above 128 repeats the next byte count-128 times, else count bytes
// follow).  A pixel (r, g, b, e) is (r, g, b) * 2^(e-136) with e = 0 black
// (G. Ward, "Real pixels", Graphics Gems II, 1991).
static bool rgbe_scanline(FILE* f, int w, unsigned char* line) {
    unsigned char head[4];
    if (std::fread(head, 1, 4, f) != 4) return false;
    if (w < 8 || w > 0x7fff || head[0] != 2 || head[1] != 2 || (head[2] & 0x80)) {  // flat
        std::memcpy(line, head, 4);
        return std::fread(line + 4, 4, (size_t)w - 1, f) == (size_t)w - 1;
    }
    if (((int)head[2] << 8 | head[3]) != w) return false;
    std::vector<unsigned char> ch((size_t)w);
    for (int c = 0; c < 4; ++c) {
        int x = 0;
        while (x < w) {
            int count = std::fgetc(f);
            if (count == EOF) return false;
            if (count > 128) {
                count -= 128;
                const int val = std::fgetc(f);
                if (val == EOF || x + count > w) return false;
                for (int k = 0; k < count; ++k) ch[x++] = (unsigned char)val;
            } else {
                if (count == 0 || x + count > w || std::fread(&ch[x], 1, count, f) != (size_t)count) return false;
                x += count;
            }
        }
        for (int i = 0; i < w; ++i) line[4 * i + c] = ch[i];
    }
    return true;
}

int mrp_read_hdr(const char* path, float* rgba_out, size_t max_texels, int* width, int* height) {
    if (!path || !width || !height) return fail(MRP_E_INVALID, "NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(MRP_E_IO, std::string("cannot open ") + path);
    char buf[512];
    bool ok = std::fgets(buf, sizeof(buf), f) && buf[0] == '#' && buf[1] == '?';
    while (ok) {
        if (!std::fgets(buf, sizeof(buf), f)) {
            ok = false;
            break;
        }
        if (buf[0] == '\n' || buf[0] == '\r') break;
        if (std::strncmp(buf, "FORMAT=", 7) == 0 && std::strncmp(buf + 7, "32-bit_rle_rgbe", 15) != 0) ok = false;
    }
    int w = 0, h = 0;
    if (ok) ok = std::fgets(buf, sizeof(buf), f) && std::sscanf(buf, "-Y %d +X %d", &h, &w) == 2 && w > 0 && h > 0 &&
                 (long long)w * h <= (1ll << 28);
    if (!ok) {
        std::fclose(f);
        return fail(MRP_E_IO, std::string("not a Radiance RGBE file with a -Y h +X w resolution: ") + path);
    }
    *width = w;
    *height = h;
    if (!rgba_out) {  // size query
        std::fclose(f);
        return MRP_OK;
    }
    if (max_texels < (size_t)w * h) {
        std::fclose(f);
        return fail(MRP_E_INVALID, "output buffer holds fewer than width * height texels");
    }
    std::vector<unsigned char> line(4 * (size_t)w);
    for (int y = 0; y < h && ok; ++y) {
        ok = rgbe_scanline(f, w, line.data());
        for (int x = 0; ok && x < w; ++x) {
            const unsigned char* p = &line[4 * (size_t)x];
            const float s = p[3] ? std::ldexp(1.0f, (int)p[3] - 136) : 0.0f;
            float* o = rgba_out + 4 * ((size_t)y * w + x);
            o[0] = (float)p[0] * s;
            o[1] = (float)p[1] * s;
            o[2] = (float)p[2] * s;
            o[3] = 1.0f;
        }
    }
    std::fclose(f);
    if (!ok) return fail(MRP_E_IO, std::string("truncated or malformed RGBE scanlines: ") + path);
    return MRP_OK;
}

int mrp_volume_destroy(mrp_volume* h) {
    if (!h) return MRP_OK;
    Volume* v = &h->v;
    if (v->stream) hipStreamSynchronize(v->stream);
    hipFree(v->grid);
    hipFree(v->cells);
    hipFree(v->quads);
    hipFree(v->occ);
    for (int m = 0; m < kNumMips; ++m) hipFree(v->mips[m]);
    for (int m = 0; m < kNumTrMips; ++m) hipFree(v->tr_mips[m]);
    hipFree(v->hg_lut);
    for (int m = 0; m < kNumMips; ++m) hipFree(v->mip_cells[m]);
    for (int m = 0; m < kNumTrMips; ++m) hipFree(v->tr_cells[m]);
    hipFree(v->lut_cells);
    hipFree(v->mip_cells16);
    hipFree(v->tr_cells16);
    hipFree(v->hdri);
    hipFree(v->queues);
    hipFree(v->counters);
    hipFree(v->samples);
    hipFree(v->nn_scratch);
    hipFree(v->serve_scratch);
    for (hipEvent_t e : v->nn_events)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : v->serve_events)
        if (e) hipEventDestroy(e);
    if (v->host_words) hipHostFree(v->host_words);
    if (v->stream) hipStreamDestroy(v->stream);
    delete h;
    return MRP_OK;
}

int mrp_volume_resolution(const mrp_volume* h, int* res) {
    if (!h || !res) return fail(MRP_E_INVALID, "NULL argument");
    *res = h->v.res;
    return MRP_OK;
}

int mrp_volume_max_density(const mrp_volume* h, float* out) {
    if (!h || !out) return fail(MRP_E_INVALID, "NULL argument");
    if (!h->v.updated) return fail(MRP_E_STATE, "mrp_volume_update() has not run");
    *out = h->v.max_density;
    return MRP_OK;
}

int mrp_set_sample_buffer_budget(mrp_volume* h, size_t bytes) {
    if (!h) return fail(MRP_E_INVALID, "NULL volume");
    if (bytes < ((size_t)1 << 20)) return fail(MRP_E_INVALID, "budget must be >= 1 MiB");
    h->v.sample_budget = bytes;
    return MRP_OK;
}

}  // extern "C"
