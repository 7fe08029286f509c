// mrp_train.hip - training of the MRPNN network on gfx950 (forward, backward,
// Adam), derived from the inference network of `RadiancePredict`
// (core/radiancePredict.cu:370-486); the reference trains a different model
// in PyTorch (train/train.py:220-254,536-546) with Adam at lr 1e-3.
//
// One step over a batch of B samples:
//   1. train_fwd_bwd_kernel - one thread per sample runs the forward and the
//      backward of mrp_train.h in FP32 and writes, for each of the 86 dense
//      layers, its input x, output y and pre-activation gradient dz into a
//      structure-of-arrays record rec[slot][sample] (coalesced across lanes);
//   2. wgrad_kernel - per layer, dW = dZ^T X and db = sum dZ over the batch:
//      16x16 output tiles, batch chunks of 64 staged through LDS, FP64
//      accumulation; the batch is split into slices of 512 samples (one
//      workgroup per tile and slice) whose partial sums wgrad_reduce_kernel
//      adds in slice order (deterministic: fixed order, no atomics);
//   3. loss_sum_kernel - FP64 sum of the per-sample losses in a fixed tree;
//   4. adam_kernel - Adam (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master
//      weights with FP64 gradients.
// Precision 1 replaces step 1 by the BF16 matrix-core forward and backward of
// mrp_train_mfma.hip (same record, same steps 2-4).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "mrp_internal.h"
#include "mrp_mlp.h"
#include "mrp_train.h"

namespace mrp {

__constant__ TrainLayer c_train_layers[kTrainLayers];

struct GpuRec {
    float* base;
    uint32_t B, s;
    __device__ __forceinline__ float get(int f) const { return base[(size_t)f * B + s]; }
    __device__ __forceinline__ void put(int f, float v) { base[(size_t)f * B + s] = v; }
};

__global__ __launch_bounds__(64) void train_fwd_bwd_kernel(const float* __restrict__ w, const float* __restrict__ feat,
                                                           const float4* __restrict__ aux, uint32_t B, float inv_batch,
                                                           float* __restrict__ rec, float* __restrict__ loss,
                                                           float* __restrict__ pred) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    GpuRec r{rec, B, s};
    const float4 x = aux[s];
    float p = 0.0f;
    loss[s] = train_sample<float>(w, c_train_layers, feat + (size_t)s * kFeatures, x.x, x.y, x.z, x.w, inv_batch, r,
                                  &p);
    pred[s] = p;
}

// One 16x16 tile of one layer's gradient [out][in + 1] (column `in` = bias).
struct WTile {
    int layer, i0, j0;
};

constexpr int kWT = 16, kWK = 64;
constexpr uint32_t kWSlice = 512;  // samples per gradient slice
constexpr int kWMaxSlices = 32;

// Slice blockIdx.y covers samples [y * slice, min(B, (y + 1) * slice)); its
// partial gradient goes to grad + y * MRP_W_TOTAL.
__global__ __launch_bounds__(256) void wgrad_kernel(const WTile* __restrict__ tiles, const float* __restrict__ rec,
                                                    uint32_t B, uint32_t slice, double* __restrict__ grad) {
    __shared__ float sdz[kWT][kWK + 1];
    __shared__ float sx[kWT][kWK + 1];
    const WTile t = tiles[blockIdx.x];
    const TrainLayer l = c_train_layers[t.layer];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // ty -> output row i, tx -> input column j
    const int i = t.i0 + ty, j = t.j0 + tx;
    const uint32_t s_begin = blockIdx.y * slice, s_end = min(B, s_begin + slice);
    grad += (size_t)blockIdx.y * MRP_W_TOTAL;
    double acc = 0.0;
    for (uint32_t s0 = s_begin; s0 < s_end; s0 += kWK) {
        // stage dz rows i0..i0+15 and x rows j0..j0+15 for samples s0..s0+63
        for (int e = threadIdx.x; e < kWT * kWK; e += 256) {
            const int r = e / kWK, c = e - r * kWK;
            const uint32_t s = s0 + (uint32_t)c;
            const int ri = t.i0 + r, rj = t.j0 + r;
            float dz = 0.0f, xv = 0.0f;
            if (s < s_end) {
                if (ri < l.out) dz = rec[(size_t)(l.dz_off + ri) * B + s];
                if (rj < l.in) xv = rec[(size_t)(l.x_off + rj) * B + s];
                else if (rj == l.in) xv = 1.0f;  // bias column
            }
            sdz[r][c] = dz;
            sx[r][c] = xv;
        }
        __syncthreads();
        for (int c = 0; c < kWK; ++c) acc += (double)sdz[ty][c] * (double)sx[tx][c];
        __syncthreads();
    }
    if (i < l.out) {
        if (j < l.in) grad[l.w_off + i * l.in + j] = acc;
        else if (j == l.in && l.b_off >= 0) grad[l.b_off + i] = acc;
    }
}

// grad[i] = sum of the slices' partial gradients, in slice order.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const double* __restrict__ part, int slices,
                                                           double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MRP_W_TOTAL) return;
    double a = 0.0;
    for (int q = 0; q < slices; ++q) a += part[(size_t)q * MRP_W_TOTAL + i];
    grad[i] = a;
}

// Sum of n values in FP64 in a fixed order: 256 strided partial sums, then a tree.
__global__ __launch_bounds__(256) void loss_sum_kernel(const float* __restrict__ v, uint32_t n,
                                                       double* __restrict__ out) {
    __shared__ double part[256];
    double a = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) a += (double)v[i];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                                   const double* __restrict__ grad, int n, float lr, float b1, float b2,
                                                   float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = (float)grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;  // bias corrections 1 - beta^t
    w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

__global__ void train_clear_kernel(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace mrp

using namespace mrp;

struct mrp_trainer {
    int device = 0;
    hipStream_t stream = nullptr;
    int max_batch = 0, rec_floats = 0;
    float* w = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    double* grad = nullptr;
    double* grad_part = nullptr;  // kWMaxSlices partial gradients
    float* rec = nullptr;
    float* loss = nullptr;
    float* pred = nullptr;
    double* loss_sum = nullptr;
    WTile* tiles = nullptr;
    int n_tiles = 0;
    int steps = 0;
    int precision = 0;          // 0: FP32 forward/backward, 1: BF16 matrix cores
    TrainMfma* mfma = nullptr;  // created on first use of precision 1
};

static void trainer_free(mrp_trainer* t) {
    if (!t) return;
    hipFree(t->w);
    hipFree(t->m);
    hipFree(t->v);
    hipFree(t->grad);
    hipFree(t->grad_part);
    hipFree(t->rec);
    hipFree(t->loss);
    hipFree(t->pred);
    hipFree(t->loss_sum);
    hipFree(t->tiles);
    train_mfma_destroy(t->mfma);
    if (t->stream) hipStreamDestroy(t->stream);
    delete t;
}

static int trainer_gradient(mrp_trainer* t, int batch, const float* feat, const float* aux) {
    if (batch <= 0 || batch > t->max_batch) return fail(MRP_E_INVALID, "batch must be in [1, max_batch]");
    if (!feat || !aux) return fail(MRP_E_INVALID, "NULL argument");
    const uint32_t B = (uint32_t)batch;
    if (t->precision == 1) {
        const int rc = train_mfma_fwd_bwd(t->mfma, t->w, feat, aux, B, 1.0f / (float)B, t->rec, t->loss, t->pred,
                                          t->loss_sum, t->stream);
        if (rc) return rc;
    } else {
        train_fwd_bwd_kernel<<<(B + 63) / 64, 64, 0, t->stream>>>(t->w, feat, (const float4*)aux, B,
                                                                  1.0f / (float)B, t->rec, t->loss, t->pred);
        MRP_HIP(hipGetLastError());
        loss_sum_kernel<<<1, 256, 0, t->stream>>>(t->loss, B, t->loss_sum);
        MRP_HIP(hipGetLastError());
    }
    const int slices = std::min<int>(kWMaxSlices, (int)((B + kWSlice - 1) / kWSlice));
    const uint32_t slice = ((B + (uint32_t)slices - 1) / (uint32_t)slices + kWK - 1) / kWK * kWK;
    wgrad_kernel<<<dim3((unsigned)t->n_tiles, (unsigned)slices), 256, 0, t->stream>>>(
        t->tiles, t->rec, B, slice, slices == 1 ? t->grad : t->grad_part);
    MRP_HIP(hipGetLastError());
    if (slices > 1) {
        wgrad_reduce_kernel<<<(MRP_W_TOTAL + 255) / 256, 256, 0, t->stream>>>(t->grad_part, slices, t->grad);
        MRP_HIP(hipGetLastError());
    }
    return MRP_OK;
}

extern "C" {

int mrp_trainer_create(const float* weights, size_t count, int max_batch, mrp_trainer** out) {
    if (!weights || !out) return fail(MRP_E_INVALID, "NULL argument");
    *out = nullptr;
    if (count != (size_t)MRP_W_TOTAL) return fail(MRP_E_INVALID, "weight count must be 49673 (NNWeight.cuh layout)");
    if (max_batch <= 0 || max_batch > (1 << 20)) return fail(MRP_E_INVALID, "max_batch must be in [1, 2^20]");
    mrp_trainer* t = new (std::nothrow) mrp_trainer();
    if (!t) return fail(MRP_E_NOMEM, "host allocation failed");
    hipGetDevice(&t->device);
    TrainLayer layers[kTrainLayers];
    t->rec_floats = build_train_layers(layers);
    t->max_batch = max_batch;
    std::vector<WTile> tiles;
    for (int li = 0; li < kTrainLayers; ++li)
        for (int i0 = 0; i0 < layers[li].out; i0 += kWT)
            for (int j0 = 0; j0 < layers[li].in + 1; j0 += kWT) tiles.push_back(WTile{li, i0, j0});
    t->n_tiles = (int)tiles.size();
    const size_t n = count;
    bool ok = hipStreamCreate(&t->stream) == hipSuccess;
    ok = ok && hipMalloc(&t->w, n * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->m, n * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->v, n * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->grad, n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&t->grad_part, (size_t)kWMaxSlices * n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&t->rec, (size_t)t->rec_floats * max_batch * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->loss, (size_t)max_batch * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->pred, (size_t)max_batch * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&t->loss_sum, sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&t->tiles, tiles.size() * sizeof(WTile)) == hipSuccess;
    if (!ok) {
        trainer_free(t);
        return fail(MRP_E_NOMEM, "trainer allocation failed");
    }
    ok = hipMemcpy(t->w, weights, n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(t->tiles, tiles.data(), tiles.size() * sizeof(WTile), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpyToSymbol(HIP_SYMBOL(c_train_layers), layers, sizeof(layers)) == hipSuccess;
    if (ok) {
        train_clear_kernel<<<(unsigned)((n + 255) / 256), 256, 0, t->stream>>>(t->m, n);
        train_clear_kernel<<<(unsigned)((n + 255) / 256), 256, 0, t->stream>>>(t->v, n);
        ok = hipStreamSynchronize(t->stream) == hipSuccess;
    }
    if (!ok) {
        trainer_free(t);
        return fail(MRP_E_HIP, "trainer initialisation failed");
    }
    *out = t;
    return MRP_OK;
}

int mrp_trainer_destroy(mrp_trainer* t) {
    trainer_free(t);
    return MRP_OK;
}

int mrp_trainer_set_precision(mrp_trainer* t, int precision) {
    if (!t) return fail(MRP_E_INVALID, "NULL trainer");
    if (precision != 0 && precision != 1) return fail(MRP_E_INVALID, "precision must be 0 (FP32) or 1 (BF16)");
    MRP_HIP(hipSetDevice(t->device));
    if (precision == 1 && !t->mfma) {
        const int rc = train_mfma_create(t->max_batch, &t->mfma);
        if (rc) return rc;
    }
    t->precision = precision;
    return MRP_OK;
}

int mrp_trainer_record_floats(const mrp_trainer* t, int* out) {
    if (!t || !out) return fail(MRP_E_INVALID, "NULL argument");
    *out = t->rec_floats;
    return MRP_OK;
}

int mrp_trainer_gradient(mrp_trainer* t, int batch, const float* feat_dev, const float* aux_dev, double* grad_host,
                         double* loss_out) {
    if (!t) return fail(MRP_E_INVALID, "NULL trainer");
    MRP_HIP(hipSetDevice(t->device));
    int rc = trainer_gradient(t, batch, feat_dev, aux_dev);
    if (rc) return rc;
    if (grad_host)
        MRP_HIP(hipMemcpyAsync(grad_host, t->grad, (size_t)MRP_W_TOTAL * sizeof(double), hipMemcpyDeviceToHost,
                               t->stream));
    double ls = 0.0;
    MRP_HIP(hipMemcpyAsync(&ls, t->loss_sum, sizeof(double), hipMemcpyDeviceToHost, t->stream));
    MRP_HIP(hipStreamSynchronize(t->stream));
    if (loss_out) *loss_out = ls / (double)batch;
    return MRP_OK;
}

int mrp_trainer_step(mrp_trainer* t, int batch, const float* feat_dev, const float* aux_dev, float lr,
                     double* loss_out) {
    if (!t) return fail(MRP_E_INVALID, "NULL trainer");
    if (!(lr > 0.0f)) return fail(MRP_E_INVALID, "lr must be > 0");
    MRP_HIP(hipSetDevice(t->device));
    int rc = trainer_gradient(t, batch, feat_dev, aux_dev);
    if (rc) return rc;
    t->steps += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - std::pow(0.9, t->steps)), c2 = (float)(1.0 - std::pow(0.999, t->steps));
    adam_kernel<<<(MRP_W_TOTAL + 255) / 256, 256, 0, t->stream>>>(t->w, t->m, t->v, t->grad, MRP_W_TOTAL, lr, b1, b2,
                                                                  c1, c2, 1e-8f);
    MRP_HIP(hipGetLastError());
    double ls = 0.0;
    MRP_HIP(hipMemcpyAsync(&ls, t->loss_sum, sizeof(double), hipMemcpyDeviceToHost, t->stream));
    MRP_HIP(hipStreamSynchronize(t->stream));
    if (loss_out) *loss_out = ls / (double)batch;
    return MRP_OK;
}

int mrp_trainer_get_weights(const mrp_trainer* t, float* host_out) {
    if (!t || !host_out) return fail(MRP_E_INVALID, "NULL argument");
    MRP_HIP(hipSetDevice(t->device));
    MRP_HIP(hipStreamSynchronize(t->stream));
    MRP_HIP(hipMemcpy(host_out, t->w, (size_t)MRP_W_TOTAL * sizeof(float), hipMemcpyDeviceToHost));
    return MRP_OK;
}

int mrp_trainer_predictions(const mrp_trainer* t, int batch, float* host_out) {
    if (!t || !host_out) return fail(MRP_E_INVALID, "NULL argument");
    if (batch <= 0 || batch > t->max_batch) return fail(MRP_E_INVALID, "batch out of range");
    MRP_HIP(hipSetDevice(t->device));
    MRP_HIP(hipStreamSynchronize(t->stream));
    MRP_HIP(hipMemcpy(host_out, t->pred, (size_t)batch * sizeof(float), hipMemcpyDeviceToHost));
    return MRP_OK;
}

}  // extern "C"
