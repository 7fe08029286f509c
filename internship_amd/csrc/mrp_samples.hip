// mrp_samples.hip - training targets of the learned part on gfx950:
// `GetSample` (core/render.cuh:276-315) as driven by `GetSamples` /
// `GetSampleMulti` (core/volume.cu:36-47,550-626).
//
// From a scattering point, a direction is drawn from the HG lobe about the
// query direction, then a delta-tracking path of up to multi_scatter bounces
// collects next-event light at every vertex (the albedo per bounce is
// 1.001 * scatter, as the reference writes it) and the sky on escape.  The
// single-scatter term at the point itself is not part of the target: the
// network predicts only the higher orders (`NNPredict` adds Tr * HG itself).
//
// Layout: one thread per (query, sample) path writes its radiance to a
// [sample][query] buffer; a second kernel adds the paths of each query in
// sample order.  Queries are processed in chunks of samples so the buffer
// stays bounded; the running sums carry over between chunks in order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mrp_internal.h"

namespace mrp {

struct SampleArgs {
    const float4* cells;
    int res;
    float max_density, tr_scale;
    vec3 light_color, sky;
    int multi_scatter;
    uint32_t seed_lo, seed_hi;
    int n;           // queries
    int s_begin;     // first sample index of this chunk
    int s_count;     // samples in this chunk
    const mrp_sample_query* q;
    float* paths;    // s_count x n x 3
    unsigned long long* counters;
};

// `DeterminateNextVertex` (core/render.cuh:138-172) on the corner-expanded cells.
__device__ bool gs_next_vertex(Rng* rng, const SampleArgs& a, const Majorant& mj, float g, vec3 pos, vec3 dir,
                               float dis, vec3* npos, vec3* ndir, uint32_t* lk) {
    float t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < 10000) {
        float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax;
        if (t > dis) return false;
        rk = rng_next(rng);
        const float d = trilinear_cells(a.cells, a.res, pos + (dir * t));
        *lk += 1u;
        if (d * mj.ratio > rk) break;
        if (d < 0.0f) t -= d;
    }
    const float e0 = rng_next(rng);
    const float e1 = rng_next(rng);
    *ndir = hg_sample_dir(e0, e1, dir, g);
    *npos = (dir * t) + pos;
    return true;
}

// `Tr` (core/render.cuh:93-136).
__device__ float gs_transmittance(Rng* rng, const SampleArgs& a, const Majorant& mj, vec3 pos, vec3 dir, float dis,
                                  uint32_t* lk) {
    float tr = 1.0f, t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < 10000 && tr > 0.5f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax_tr;
        if (t > dis) break;
        const float d = trilinear_cells(a.cells, a.res, pos + (dir * t));
        *lk += 1u;
        tr *= 1.0f - fmax_(0.0f, d * mj.ratio);
        if (d < 0.0f) t -= d;
    }
    while (loop_num++ < 10000 && tr > 0.001f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax_tr;
        if (t > dis) break;
        const float d = trilinear_cells(a.cells, a.res, pos + (dir * t));
        *lk += 1u;
        if (rng_next(rng) < fmax_(0.0f, d * mj.ratio)) {
            tr = 0.0f;
            break;
        }
        if (d < 0.0f) t -= d;
    }
    return tr;
}

__global__ __launch_bounds__(256) void get_sample_kernel(SampleArgs a) {
    __shared__ unsigned red[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lk = 0u;
    if (i < (uint32_t)a.n * (uint32_t)a.s_count) {
        const uint32_t k = i / (uint32_t)a.n, j = i - k * (uint32_t)a.n;
        const mrp_sample_query q = a.q[j];
        const Majorant mj = make_majorant(a.max_density, q.alpha, a.tr_scale);
        const vec3 light = v3(q.light_dir.x, q.light_dir.y, q.light_dir.z);
        Rng rng = rng_make(j, (uint32_t)(a.s_begin + (int)k), 1u, ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        vec3 pos = v3(q.pos.x, q.pos.y, q.pos.z);
        const float e0 = rng_next(&rng);
        const float e1 = rng_next(&rng);
        vec3 ray = hg_sample_dir(e0, e1, v3(q.dir.x, q.dir.y, q.dir.z), q.g);
        vec3 csr = v3(1, 1, 1), res = v3(0, 0, 0);
        for (int s = 0; s < a.multi_scatter; ++s) {
            vec3 np, nd;
            if (!gs_next_vertex(&rng, a, mj, q.g, pos, ray, ray_box_distance(pos, ray), &np, &nd, &lk)) {
                res = res + a.sky * csr;
                break;
            }
            pos = np;
            csr = (csr * 1.001f) * q.scatter;
            const float phase = hg_phase(dot(ray, light), q.g);
            const float tr = gs_transmittance(&rng, a, mj, pos, light, ray_box_distance(pos, light), &lk);
            res = res + a.light_color * csr * (tr * phase);
            ray = nd;
        }
        float* o = a.paths + 3 * (size_t)i;
        o[0] = res.x;
        o[1] = res.y;
        o[2] = res.z;
    }
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&a.counters[0], (unsigned long long)red[0]);
}

// sums[j] += paths[k][j] for k = 0..count-1 in order; on the last chunk
// out[j] = sums[j] / spp_total (`res / sampleNum`).
__global__ __launch_bounds__(256) void sample_sum_kernel(const float* __restrict__ paths, int n, int count,
                                                         float* __restrict__ sums, int spp_total, int last,
                                                         float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    vec3 s = v3(sums[3 * j], sums[3 * j + 1], sums[3 * j + 2]);
    for (int k = 0; k < count; ++k) {
        const float* p = paths + 3 * ((size_t)k * n + j);
        s = s + v3(p[0], p[1], p[2]);
    }
    sums[3 * j] = s.x;
    sums[3 * j + 1] = s.y;
    sums[3 * j + 2] = s.z;
    if (last) {
        const float d = (float)spp_total;
        out[3 * j] = s.x / d;
        out[3 * j + 1] = s.y / d;
        out[3 * j + 2] = s.z / d;
    }
}

__global__ void sample_clear_kernel(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_get_samples(mrp_volume* vol, int n, const mrp_sample_query* q_dev, mrp_vec3 light_color,
                               mrp_vec3 sky, int multi_scatter, int spp, uint64_t seed, float* out_dev,
                               mrp_trace_stats* stats) {
    if (!vol || (n > 0 && (!q_dev || !out_dev))) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0 || spp <= 0 || multi_scatter < 0 || spp > (1 << 24)) return fail(MRP_E_INVALID, "bad n / spp / multi");
    if (n == 0) return MRP_OK;
    Volume* v = &vol->v;
    MRP_HIP(hipSetDevice(v->device));
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() must run before sampling");
    // a chunk of samples per launch: at most 2^26 paths (768 MiB of path radiance)
    // and no more than the volume's sample-buffer budget
    const long long max_paths = std::min<long long>((long long)1 << 26, (long long)(v->sample_budget / 12));
    const int chunk = (int)std::max<long long>(1, std::min<long long>(spp, max_paths / n));
    const size_t path_bytes = (size_t)chunk * n * 3 * sizeof(float);
    int rc = ensure_samples(v, path_bytes + (size_t)n * 3 * sizeof(float));
    if (rc) return rc;
    float* paths = (float*)v->samples;
    float* sums = paths + (size_t)chunk * n * 3;
    SampleArgs a;
    a.cells = v->cells;
    a.res = v->res;
    a.max_density = v->max_density;
    a.tr_scale = 1.0f;
    a.light_color = v3(light_color.x, light_color.y, light_color.z);
    a.sky = v3(sky.x, sky.y, sky.z);
    a.multi_scatter = multi_scatter;
    a.seed_lo = (uint32_t)seed;
    a.seed_hi = (uint32_t)(seed >> 32);
    a.n = n;
    a.q = q_dev;
    a.paths = paths;
    a.counters = v->counters;
    hipEvent_t e0, e1;
    MRP_HIP(hipEventCreate(&e0));
    MRP_HIP(hipEventCreate(&e1));
    MRP_HIP(hipEventRecord(e0, v->stream));
    rc = clear_counters(v);
    if (rc == MRP_OK) {
        sample_clear_kernel<<<(unsigned)((3 * (size_t)n + 255) / 256), 256, 0, v->stream>>>(sums, 3 * (size_t)n);
        for (int s0 = 0; s0 < spp; s0 += chunk) {
            const int cnt = std::min(chunk, spp - s0);
            a.s_begin = s0;
            a.s_count = cnt;
            const size_t items = (size_t)cnt * n;
            get_sample_kernel<<<(unsigned)((items + 255) / 256), 256, 0, v->stream>>>(a);
            MRP_HIP(hipGetLastError());
            sample_sum_kernel<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(paths, n, cnt, sums, spp,
                                                                                 s0 + cnt >= spp ? 1 : 0, out_dev);
            MRP_HIP(hipGetLastError());
        }
        MRP_HIP(hipEventRecord(e1, v->stream));
        MRP_HIP(hipEventSynchronize(e1));
    }
    float ms = 0.0f;
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    if (rc == MRP_OK) {
        MRP_HIP(hipEventElapsedTime(&ms, e0, e1));
        rc = read_counters(v, cnt4);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    if (stats) {
        *stats = mrp_trace_stats{};
        stats->lookups = cnt4[0];
        stats->paths = (uint64_t)n * (uint64_t)spp;
        stats->kernel_ms = ms;
        stats->total_ms = ms;
    }
    return MRP_OK;
}
