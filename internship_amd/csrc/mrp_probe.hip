// mrp_probe.hip - the random-read ceiling of the GPU, the bound that the
// lookup kernels (tracer, first vertex, shadow rays) are read against.
//
// A trilinear lookup reads one 32-byte cell at a data-dependent address, so an
// L2 miss moves a whole 128-byte line.  How many such random reads per second
// the memory system serves is a property of the chip, not of the tracer: this
// kernel measures it with the same access shape and no arithmetic, every
// thread reading `iters` items of `width` bytes at hashed positions in a
// buffer of `buffer_bytes` (a power of two; indices are masked to it).
#include <hip/hip_runtime.h>

#include "mrp_internal.h"

namespace mrp {

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <int W>
__global__ __launch_bounds__(256) void random_read_kernel(const uint4* __restrict__ buf, uint32_t mask, int iters,
                                                          float* __restrict__ sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    uint32_t h = probe_hash(t * 2654435761u + 12345u);
    for (int i = 0; i < iters; ++i) {
        h = probe_hash(h + (uint32_t)i);
        const size_t idx = h & mask;  // item index < buffer_bytes / W
        uint32_t x;
        if (W == 16) {
            x = buf[idx].x;
        } else if (W == 32) {
            const uint4 a = buf[2 * idx], b = buf[2 * idx + 1];
            x = a.x ^ b.w;
        } else {
            const uint4* p = buf + 8 * idx;
            x = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) x ^= p[k].y;
        }
        acc += __uint_as_float(x & 0x3f7fffffu);
    }
    if (acc == -1.0f) sink[0] = acc;  // never true for the zero-filled buffer; keeps the loads
}

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_measure_random_reads(int width, uint64_t buffer_bytes, int iters, double* reads_per_s) {
    if (!reads_per_s) return fail(MRP_E_INVALID, "NULL argument");
    if (width != 16 && width != 32 && width != 128) return fail(MRP_E_INVALID, "width must be 16, 32 or 128 bytes");
    if (buffer_bytes < ((uint64_t)1 << 20) || buffer_bytes > ((uint64_t)1 << 34) ||
        (buffer_bytes & (buffer_bytes - 1)) != 0)
        return fail(MRP_E_INVALID, "buffer_bytes must be a power of two from 1 MiB to 16 GiB");
    if (iters < 1 || iters > 4096) return fail(MRP_E_INVALID, "iters must be in [1, 4096]");
    const uint64_t items = buffer_bytes / (uint64_t)width;
    if (items > ((uint64_t)1 << 32)) return fail(MRP_E_INVALID, "too many items for 32-bit hashing");
    const uint32_t mask = (uint32_t)(items - 1u);
    void* buf = nullptr;
    float* sink = nullptr;
    if (hipMalloc(&buf, buffer_bytes) != hipSuccess) return fail(MRP_E_NOMEM, "probe buffer allocation failed");
    if (hipMalloc((void**)&sink, sizeof(float)) != hipSuccess) {
        hipFree(buf);
        return fail(MRP_E_NOMEM, "probe allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemset(buf, 0, buffer_bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const unsigned blocks = 256u * 64u;  // 4.2M threads
    float best = 0.0f;
    for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {  // the first launch warms the TLB and caches
        e = hipEventRecord(e0, nullptr);
        if (width == 16)
            random_read_kernel<16><<<blocks, 256>>>((const uint4*)buf, mask, iters, sink);
        else if (width == 32)
            random_read_kernel<32><<<blocks, 256>>>((const uint4*)buf, mask, iters, sink);
        else
            random_read_kernel<128><<<blocks, 256>>>((const uint4*)buf, mask, iters, sink);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && rep > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(buf);
    hipFree(sink);
    if (e != hipSuccess) return hip_fail(e, "random read probe");
    *reads_per_s = (double)blocks * 256.0 * (double)iters / ((double)best * 1e-3);
    return MRP_OK;
}
