// mrp_api.hip - C ABI: errors, device-memory plumbing, volume lifetime.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
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

int mrp_volume_destroy(mrp_volume* h) {
    if (!h) return MRP_OK;
    Volume* v = &h->v;
    if (v->stream) hipStreamSynchronize(v->stream);
    hipFree(v->grid);
    hipFree(v->cells);
    for (int m = 0; m < kNumMips; ++m) hipFree(v->mips[m]);
    for (int m = 0; m < kNumTrMips; ++m) hipFree(v->tr_mips[m]);
    hipFree(v->hg_lut);
    for (int m = 0; m < kNumMips; ++m) hipFree(v->mip_cells[m]);
    for (int m = 0; m < kNumTrMips; ++m) hipFree(v->tr_cells[m]);
    hipFree(v->lut_cells);
    hipFree(v->mip_cells16);
    hipFree(v->tr_cells16);
    hipFree(v->queues);
    hipFree(v->counters);
    hipFree(v->samples);
    hipFree(v->nn_scratch);
    for (hipEvent_t e : v->nn_events)
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
