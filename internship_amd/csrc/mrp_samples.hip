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
#include "mrp_track.h"

namespace mrp {

struct SampleArgs {
    const float4* cells;
    Occupancy occ;  // brick occupancy of the cells
    int res;
    float max_density, tr_scale;
    vec3 light_color;
    EnvMap env;  // `SkyBox` of escaping paths
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
__device__ bool gs_next_vertex(Rng* rng, const float4* __restrict__ cells, int res, const Majorant& mj, float g,
                               vec3 pos, vec3 dir, float dis, vec3* npos, vec3* ndir, uint32_t* lk) {
    float t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < 10000) {
        float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax;
        if (t > dis) return false;
        rk = rng_next(rng);
        const float d = trilinear_cells(cells, res, pos + (dir * t));
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
__device__ float gs_transmittance(Rng* rng, const float4* __restrict__ cells, int res, const Majorant& mj, vec3 pos,
                                  vec3 dir, float dis, uint32_t* lk) {
    float tr = 1.0f, t = 0.0f;
    int loop_num = 0;
    while (loop_num++ < 10000 && tr > 0.5f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax_tr;
        if (t > dis) break;
        const float d = trilinear_cells(cells, res, pos + (dir * t));
        *lk += 1u;
        tr *= 1.0f - fmax_(0.0f, d * mj.ratio);
        if (d < 0.0f) t -= d;
    }
    while (loop_num++ < 10000 && tr > 0.001f) {
        const float rk = rng_next(rng);
        t -= logf_(1.0f - rk) * mj.inv_smax_tr;
        if (t > dis) break;
        const float d = trilinear_cells(cells, res, pos + (dir * t));
        *lk += 1u;
        if (rng_next(rng) < fmax_(0.0f, d * mj.ratio)) {
            tr = 0.0f;
            break;
        }
        if (d < 0.0f) t -= d;
    }
    return tr;
}

// `GetSample` paths, item i = (sample k = i / n, query j = i % n) on Philox
// stream (j, s_begin + k, tag 1, seed); paths[i] = the path's radiance.
// Persistent, with lanes that refill: a lane runs one bounce of its path per
// iteration (free flight with kSampleBatch lookups in flight, then the shadow
// ray), and when at least kSampleRefill lanes of a wave are idle the wave
// claims that many items with one atomic.  The draws, arithmetic and lookups
// of each path are those of the one-lookup-at-a-time loop (mrp_track.h);
// only which lane runs it, and when, differ.  Every wave leaves once the
// items are claimed and its lanes are idle.
constexpr int kTrackSteps = 10000;  // loop tests per walk (the reference's cap)
constexpr int kSampleRefill = 16;
constexpr int kSampleBatch = 4;
constexpr int kSampleBlocksPerCU = 3;  // 3 waves per SIMD at <= 168 VGPRs
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void get_sample_kernel(SampleArgs a,
                                                                                                uint32_t* cursor) {
    __shared__ unsigned red[256];
    __shared__ uint32_t s_ring[256 * kRingWords];  // the threads' Philox rings (16 KB)
    extern __shared__ uint32_t s_occ[];             // occupancy_words(a.res) words
    const Occupancy occ = stage_occupancy(a.occ, a.res, s_occ);
    const uint32_t lane = __lane_id();
    const uint32_t n_items = (uint32_t)a.n * (uint32_t)a.s_count;
    uint32_t lk = 0u;
    bool active = false, queue_done = false;
    uint32_t i = 0u;
    int bounce = 0;
    uint32_t* const ring = lds_ring(s_ring);
    RingRng<64> rng = ring_make<64>(ring, 0u, 0u, 0u, 0ull);
    vec3 pos = v3(0, 0, 0), ray = v3(0, 0, 1), csr = v3(1, 1, 1), res = v3(0, 0, 0), light = v3(0, 1, 0);
    float g = 0.0f, scatter = 1.0f;
    Majorant mj = make_majorant(1.0f, 1.0f, 1.0f);
    const CellsFetch F{a.cells, a.res, occ};
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!queue_done && n_idle >= (uint32_t)kSampleRefill) {  // uniform over the wave
            uint32_t base = 0u;
            if (lane == 0u) base = atomicAdd(cursor, n_idle);
            base = __shfl(base, 0);
            if (base >= n_items) queue_done = true;
            if (!active) {
                const uint32_t item = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (item < n_items) {
                    i = item;
                    const uint32_t k = i / (uint32_t)a.n, j = i - k * (uint32_t)a.n;
                    const mrp_sample_query q = a.q[j];
                    mj = make_majorant(a.max_density, q.alpha, a.tr_scale);
                    light = v3(q.light_dir.x, q.light_dir.y, q.light_dir.z);
                    g = q.g;
                    scatter = q.scatter;
                    rng = ring_make<64>(ring, j, (uint32_t)(a.s_begin + (int)k), 1u,
                                        ((uint64_t)a.seed_hi << 32) | a.seed_lo);
                    pos = v3(q.pos.x, q.pos.y, q.pos.z);
                    const float e0 = rng_next(&rng);
                    const float e1 = rng_next(&rng);
                    ray = hg_sample_dir(e0, e1, v3(q.dir.x, q.dir.y, q.dir.z), q.g);
                    csr = v3(1, 1, 1);
                    res = v3(0, 0, 0);
                    bounce = 0;
                    active = a.multi_scatter > 0;
                    if (!active) {
                        float* o = a.paths + 3 * (size_t)i;
                        o[0] = o[1] = o[2] = 0.0f;
                    }
                }
            }
        }
        if (queue_done && !__any(active)) break;  // uniform: every wave reaches it once its lanes are idle
        if (active) {
            vec3 np;
            bool done = false;
            if (!free_flight_batched<kSampleBatch>(&rng, F, mj.inv_smax, mj.ratio, pos, ray, ray_box_distance(pos, ray),
                                                   kTrackSteps, &np, &lk)) {
                res = res + env_radiance(a.env, ray) * csr;
                done = true;
            } else {
                rng_seek(&rng, rng.n - 2u);  // the two direction draws the free flight consumed
                const float e0 = rng_next(&rng);
                const float e1 = rng_next(&rng);
                const vec3 nd = hg_sample_dir(e0, e1, ray, g);
                pos = np;
                csr = (csr * 1.001f) * scatter;
                const float phase = hg_phase(dot(ray, light), g);
                const float tr = transmittance_batched<kSampleBatch>(&rng, F, mj.inv_smax_tr, mj.ratio, pos, light,
                                                                     ray_box_distance(pos, light), kTrackSteps, &lk);
                res = res + a.light_color * csr * (tr * phase);
                ray = nd;
                done = ++bounce >= a.multi_scatter;
            }
            if (done) {
                float* o = a.paths + 3 * (size_t)i;
                o[0] = res.x;
                o[1] = res.y;
                o[2] = res.z;
                active = false;
            }
        }
    }
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&a.counters[0], (unsigned long long)red[0]);
}

__global__ void clear_cursor_kernel(uint32_t* c) {
    if (threadIdx.x == 0) *c = 0u;
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

// Training points (`MeanFreePathSample`, dataGen.cu:174-216) of n_rays random
// rays, one thread per ray: slots 4r..4r+3 hold the kept vertices of ray r
// (pos.w = 1 when kept, 0 otherwise; view.w = 0).  Ray r draws from Philox
// stream (r, 0, tag 2, seed).
__global__ __launch_bounds__(256) void training_points_kernel(const float4* __restrict__ cells, int res,
                                                              Majorant mj, float g, int n_rays, uint32_t seed_lo,
                                                              uint32_t seed_hi, float4* __restrict__ pos_out,
                                                              float4* __restrict__ view_out,
                                                              unsigned long long* __restrict__ counters) {
    __shared__ unsigned red[256];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lk = 0u;
    if (r < (uint32_t)n_rays) {
        Rng rng = rng_make(r, 0u, 2u, ((uint64_t)seed_hi << 32) | seed_lo);
        const auto nv = [&](Rng* rg, vec3 p, vec3 d, float dis, vec3* np, vec3* nd) {
            return gs_next_vertex(rg, cells, res, mj, g, p, d, dis, np, nd, &lk);
        };
        TrainPoint tp[4];
        training_ray_points(&rng, nv, tp);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            pos_out[4 * (size_t)r + i] = make_float4(tp[i].pos.x, tp[i].pos.y, tp[i].pos.z, tp[i].kept ? 1.0f : 0.0f);
            view_out[4 * (size_t)r + i] = make_float4(tp[i].view.x, tp[i].view.y, tp[i].view.z, 0.0f);
        }
    }
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&counters[0], (unsigned long long)red[0]);
}

// Lookup counts of a workgroup into counters[0] (one atomic per workgroup).
__device__ __forceinline__ void add_lookups(unsigned* red, uint32_t lk, unsigned long long* counters) {
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&counters[0], (unsigned long long)red[0]);
}

struct RayArgs {
    const float4* cells;
    int res;
    Majorant mj;
    float g;
    vec3 light, light_color, scatter_rate;
    EnvMap env;
    int multi_scatter, spp, n;
    uint32_t seed_lo, seed_hi;
    const float* ori;  // n x 3
    const float* dir;  // n x 3
    unsigned long long* counters;
};

// `CalculateRadianceMulti<PT>` (core/volume.cu:17-34): one thread per given
// ray, spp paths on Philox stream (ray, 0, tag 4, seed); out[i] = (mean
// radiance, first-scattering distance or -1).
__global__ __launch_bounds__(256) void radiance_rays_kernel(RayArgs a, float4* __restrict__ out) {
    __shared__ unsigned red[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lk = 0u;
    if (i < (uint32_t)a.n) {
        Rng rng = rng_make(i, 0u, 4u, ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        const auto nv = [&](Rng* rg, vec3 p, vec3 d, float dis, vec3* np, vec3* nd) {
            return gs_next_vertex(rg, a.cells, a.res, a.mj, a.g, p, d, dis, np, nd, &lk);
        };
        const auto tr = [&](Rng* rg, vec3 p, vec3 d, float dis) {
            return gs_transmittance(rg, a.cells, a.res, a.mj, p, d, dis, &lk);
        };
        const vec3 o = v3(a.ori[3 * (size_t)i], a.ori[3 * (size_t)i + 1], a.ori[3 * (size_t)i + 2]);
        const vec3 d = v3(a.dir[3 * (size_t)i], a.dir[3 * (size_t)i + 1], a.dir[3 * (size_t)i + 2]);
        float t;
        const vec3 r = radiance_paths(&rng, nv, tr, o, d, a.light, a.light_color, a.env, a.scatter_rate, a.g,
                                      a.multi_scatter, a.spp, &t);
        out[i] = make_float4(r.x, r.y, r.z, t);
    }
    add_lookups(red, lk, a.counters);
}

// `GetTrMulti` / `CalculateShadowTerm_TR` (core/volume.cu:297-300,630-636):
// one thread per point, spp estimates of Tr * HG on Philox stream (point, 0,
// tag 5, seed).
__global__ __launch_bounds__(256) void shadow_terms_kernel(RayArgs a, float* __restrict__ out) {
    __shared__ unsigned red[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t lk = 0u;
    if (i < (uint32_t)a.n) {
        Rng rng = rng_make(i, 0u, 5u, ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        const auto tr = [&](Rng* rg, vec3 p, vec3 d, float dis) {
            return gs_transmittance(rg, a.cells, a.res, a.mj, p, d, dis, &lk);
        };
        const vec3 o = v3(a.ori[3 * (size_t)i], a.ori[3 * (size_t)i + 1], a.ori[3 * (size_t)i + 2]);
        const vec3 d = v3(a.dir[3 * (size_t)i], a.dir[3 * (size_t)i + 1], a.dir[3 * (size_t)i + 2]);
        out[i] = shadow_term_paths(&rng, tr, o, d, a.light, a.g, a.spp);
    }
    add_lookups(red, lk, a.counters);
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
    a.occ = Occupancy{v->occ, occupancy_shift(v->res), occupancy_n(v->res)};
    a.res = v->res;
    a.max_density = v->max_density;
    a.tr_scale = 1.0f;
    a.light_color = v3(light_color.x, light_color.y, light_color.z);
    a.env = env_map(v, v3(sky.x, sky.y, sky.z));
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
            clear_cursor_kernel<<<1, 64, 0, v->stream>>>(v->queues);  // the tracer's queue words are free here
            const unsigned blocks = (unsigned)std::min<size_t>((items + 255) / 256,
                                                               (size_t)std::max(1, v->num_cus * kSampleBlocksPerCU));
            get_sample_kernel<<<blocks, 256, occupancy_words(a.res) * sizeof(uint32_t), v->stream>>>(a, v->queues);
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

extern "C" int mrp_sample_training_points(mrp_volume* vol, int n_rays, float alpha, float g, uint64_t seed,
                                          float* pos_dev, float* view_dev, uint64_t* lookups) {
    if (!vol || (n_rays > 0 && (!pos_dev || !view_dev))) return fail(MRP_E_INVALID, "NULL argument");
    if (n_rays < 0 || n_rays > (1 << 28)) return fail(MRP_E_INVALID, "n_rays must be in [0, 2^28]");
    if (!(alpha > 0.0f)) return fail(MRP_E_INVALID, "alpha must be > 0");
    Volume* v = &vol->v;
    MRP_HIP(hipSetDevice(v->device));
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() must run before sampling");
    if (n_rays == 0) {
        if (lookups) *lookups = 0;
        return MRP_OK;
    }
    int rc = clear_counters(v);
    if (rc) return rc;
    const Majorant mj = make_majorant(v->max_density, alpha, 1.0f);
    training_points_kernel<<<(unsigned)((n_rays + 255) / 256), 256, 0, v->stream>>>(
        v->cells, v->res, mj, g, n_rays, (uint32_t)seed, (uint32_t)(seed >> 32), (float4*)pos_dev, (float4*)view_dev,
        v->counters);
    MRP_HIP(hipGetLastError());
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    rc = read_counters(v, cnt4);
    if (rc) return rc;
    if (lookups) *lookups = cnt4[0];
    return MRP_OK;
}

static int ray_args(Volume* v, int n, const float* ori, const float* dir, vec3 light, float alpha, float tr_scale,
                    float g, int spp, uint64_t seed, RayArgs* a) {
    if (!ori || !dir) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0 || n > (1 << 28)) return fail(MRP_E_INVALID, "n must be in [0, 2^28]");
    if (spp <= 0 || spp > (1 << 24)) return fail(MRP_E_INVALID, "spp must be in [1, 2^24]");
    if (!(alpha > 0.0f)) return fail(MRP_E_INVALID, "alpha must be > 0");
    MRP_HIP(hipSetDevice(v->device));
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() must run first");
    a->cells = v->cells;
    a->res = v->res;
    a->mj = make_majorant(v->max_density, alpha, tr_scale);
    a->g = g;
    a->light = normalize(light);
    a->spp = spp;
    a->n = n;
    a->seed_lo = (uint32_t)seed;
    a->seed_hi = (uint32_t)(seed >> 32);
    a->ori = ori;
    a->dir = dir;
    a->counters = v->counters;
    return MRP_OK;
}

extern "C" int mrp_get_radiances(mrp_volume* vol, int n, const float* ori_dev, const float* dir_dev,
                                 const mrp_render_params* p, float* out_dev, mrp_trace_stats* stats) {
    if (!vol || !p || (n > 0 && !out_dev)) return fail(MRP_E_INVALID, "NULL argument");
    Volume* v = &vol->v;
    if (p->render_type == MRP_RENDER_MRPNN) return mrpnn_rays(v, n, ori_dev, dir_dev, p, out_dev, stats);
    if (p->render_type != MRP_RENDER_PT) return fail(MRP_E_INVALID, "render_type must be PT or MRPNN");
    if (p->multi_scatter < 0) return fail(MRP_E_INVALID, "multi_scatter must be >= 0");
    RayArgs a;
    int rc = ray_args(v, n, ori_dev, dir_dev, v3(p->light_dir.x, p->light_dir.y, p->light_dir.z), p->alpha,
                      p->tr_scale, p->g, p->spp_count, p->seed, &a);
    if (rc || n == 0) return rc;
    a.light_color = v3(p->light_color.x, p->light_color.y, p->light_color.z);
    a.env = env_map(v, v3(p->sky.x, p->sky.y, p->sky.z));
    a.scatter_rate = v3(p->scatter_rate.x, p->scatter_rate.y, p->scatter_rate.z);
    a.multi_scatter = p->multi_scatter;
    rc = clear_counters(v);
    if (rc) return rc;
    radiance_rays_kernel<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(a, (float4*)out_dev);
    MRP_HIP(hipGetLastError());
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    rc = read_counters(v, cnt4);
    if (rc) return rc;
    if (stats) {
        *stats = mrp_trace_stats{};
        stats->lookups = cnt4[0];
        stats->paths = (uint64_t)n * (uint64_t)p->spp_count;
    }
    return MRP_OK;
}

extern "C" int mrp_get_shadow_terms(mrp_volume* vol, int n, const float* pos_dev, const float* dir_dev,
                                    mrp_vec3 light_dir, float alpha, float g, int spp, uint64_t seed, float* out_dev,
                                    uint64_t* lookups) {
    if (!vol || (n > 0 && !out_dev)) return fail(MRP_E_INVALID, "NULL argument");
    Volume* v = &vol->v;
    RayArgs a;
    int rc = ray_args(v, n, pos_dev, dir_dev, v3(light_dir.x, light_dir.y, light_dir.z), alpha, 1.0f, g, spp, seed,
                      &a);
    if (rc) return rc;
    if (n == 0) {
        if (lookups) *lookups = 0;
        return MRP_OK;
    }
    rc = clear_counters(v);
    if (rc) return rc;
    shadow_terms_kernel<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(a, out_dev);
    MRP_HIP(hipGetLastError());
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    rc = read_counters(v, cnt4);
    if (rc) return rc;
    if (lookups) *lookups = cnt4[0];
    return MRP_OK;
}
