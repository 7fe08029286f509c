// mrp_serve.hip - batched serving of several scenes and cameras on gfx950.
//
// One call renders n views (volume, camera, light) of one image size:
//   1. the MRPNN stages over every view at once (mrpnn_views: one launch per
//      stage for the whole batch, items view-major), mrpnn_spp samples per
//      pixel, exactly the samples mrp_render_offline's MRPNN path draws;
//   2. serve_accumulate_kernel folds each pixel's samples into the running
//      mean and the per-pixel histogram of the frame path (`RenderCamera`,
//      core/volume.cu:99-123) and lists the pixels whose histogram variance
//      |E[l^2] - E[l]^2| - the statistic `Denoise` thresholds
//      (core/volume.cu:229-231) - exceeds the caller's threshold;
//   3. fallback_kernel re-estimates the listed pixels of all views with
//      delta tracking (`CalculateRadiance`, core/volume.cu:17-34): one
//      persistent launch (per 16M-sample pass) whose lanes take (pixel,
//      sample) paths from one queue as their previous path ends, so scenes
//      of different density and size balance across the chip; the walks keep
//      several lookups in flight per lane (mrp_track.h);
//      fallback_finish_kernel then folds each pixel's samples, in order,
//      into its mean and histogram;
//   4. the frame path's denoise and BGRA8 pack over every view (one launch,
//      blockIdx.z = view).
// The reference renders one scene per call and has no such mode; the parts
// are its algorithms, composed for a render service.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mrp_frame.h"
#include "mrp_internal.h"
#include "mrp_track.h"

namespace mrp {

constexpr int kFallbackBatch = 4;      // lookups in flight per lane
constexpr int kFallbackSteps = 10000;  // loop tests per walk (the reference's cap)
constexpr int kFallbackBlocksPerCU = 3;  // 3 waves per SIMD at <= 168 VGPRs
constexpr size_t kFallbackPassItems = (size_t)1 << 24;  // 256 MiB of per-sample results per pass

// What a fallback path needs of its view.
struct FallbackView {
    const float4* cells;
    int res;
    Majorant mj;
    float g;
    vec3 ori, forward, right, up, light, light_color, scatter_rate;
    EnvMap env;
    int width, height, multi_scatter;
    uint32_t seed_lo, seed_hi;
};

struct ServeArgs {
    uint32_t n_views, P;
    int spp, fallback_spp;
    float threshold;
};

__global__ void clear_u32_pair_kernel(uint32_t* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0u;
}

// One thread per (view, pixel): mean and histogram of the MRPNN samples,
// the fallback flag, and the list of flagged pixels (a wave ballot and one
// atomic per wave; the list order depends on scheduling, no result does).
// The block's histograms are built in LDS (a stride of 33 floats: no bank
// conflicts) and stored as one contiguous run of 256 x 33 floats.
constexpr int kAccBlock = 256;
__global__ __launch_bounds__(kAccBlock) void serve_accumulate_kernel(ServeArgs a, const float4* __restrict__ samples,
                                                                     float* __restrict__ rgb, float* __restrict__ histo,
                                                                     uint8_t* __restrict__ flags,
                                                                     uint32_t* __restrict__ list,
                                                                     uint32_t* __restrict__ count) {
    __shared__ float s_h[kAccBlock * kHistoFloats];
    const size_t total = (size_t)a.n_views * a.P;
    const size_t g0 = (size_t)blockIdx.x * kAccBlock;
    const size_t gi = g0 + threadIdx.x;
    bool flag = false;
    float* hp = s_h + kHistoFloats * threadIdx.x;
    if (gi < total) {
        const size_t view = gi / a.P, pix = gi - view * a.P;
        float tg[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < a.spp; ++s) {
            const float4 r = samples[(view * (size_t)a.spp + (size_t)s) * a.P + pix];
            histo_accumulate(tg, hp, vmax(v3(0, 0, 0), v3(r.x, r.y, r.z)), s);
        }
        rgb[3 * gi] = tg[0];
        rgb[3 * gi + 1] = tg[1];
        rgb[3 * gi + 2] = tg[2];
        const float x = hp[3 * kHistoBins + 2], x2 = hp[3 * kHistoBins + 1];
        flag = fabs_(x2 - x * x) > a.threshold;
        if (flags) flags[gi] = flag ? 1 : 0;
    }
    const unsigned long long m = __ballot(flag);
    const uint32_t lane = __lane_id();
    uint32_t base = 0u;
    if (lane == 0u && m) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __builtin_amdgcn_readfirstlane(base);  // lane 0 is active: no thread has left
    if (flag) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)gi;
    __syncthreads();
    const size_t n_here = total - g0 < (size_t)kAccBlock ? total - g0 : (size_t)kAccBlock;
    float* out = histo + (size_t)kHistoFloats * g0;
    for (size_t i = threadIdx.x; i < n_here * kHistoFloats; i += kAccBlock) out[i] = s_h[i];
}

// The fallback paths.  Item it of the pass is sample k = it % fallback_spp of
// listed pixel j = j0 + it / fallback_spp: sample spp + k on Philox stream
// (pixel, spp + k, 0, seed), a `CalculateRadiance` path, stored at out[it];
// lookups go to counters[3].  Persistent, with lanes that refill: a lane runs
// one bounce of its path per iteration (free flight, then the shadow ray),
// and when at least kFallbackRefill lanes of a wave have finished their
// paths the wave claims that many items with one atomic and starts them, so
// paths of up to 512 bounces and paths that escape at once share waves
// without idling them (25.6 -> 17.0 ms against one path per lane).  Each
// path is `radiance_paths` with spp = 1 step by step (same draws, same
// arithmetic, same lookups); only which lane runs it, and when, differ.
// Every wave leaves once the pass's items are claimed and its lanes are idle.
constexpr int kFallbackRefill = 16;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void fallback_kernel(ServeArgs a, const FallbackView* __restrict__ views,
                                                              const uint32_t* __restrict__ list, uint32_t j0,
                                                              uint32_t n_items, uint32_t* __restrict__ cursor,
                                                              float4* __restrict__ out,
                                                              unsigned long long* __restrict__ counters) {
    __shared__ unsigned red[256];
    __shared__ uint32_t s_ring[256 * kRingWords];  // the threads' Philox rings (16 KB)
    const uint32_t lane = __lane_id();
    uint32_t lk = 0u;
    bool active = false, queue_done = false;
    uint32_t it = 0u, view = 0u;
    int bounce = 0;
    uint32_t* const ring = lds_ring(s_ring);
    RingRng<64> rng = ring_make<64>(ring, 0u, 0u, 0u, 0ull);
    vec3 ori = v3(0, 0, 0), pos = v3(0, 0, 0), dir = v3(0, 0, 1), csr = v3(1, 1, 1), res = v3(0, 0, 0);
    float t_first = -1.0f;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!queue_done && n_idle >= (uint32_t)kFallbackRefill) {  // uniform over the wave
            uint32_t base = 0u;
            if (lane == 0u) base = atomicAdd(cursor, n_idle);
            base = __shfl(base, 0);
            if (base >= n_items) queue_done = true;
            if (!active) {
                const uint32_t item = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (item < n_items) {
                    it = item;
                    const uint32_t jj = it / (uint32_t)a.fallback_spp, k = it - jj * (uint32_t)a.fallback_spp;
                    const uint32_t gi = list[j0 + jj];
                    view = gi / a.P;
                    const uint32_t pix = gi - view * a.P;
                    const FallbackView& v = views[view];
                    const int row = (int)pix / v.width, col = (int)pix - row * v.width;
                    rng = ring_make<64>(ring, pix, (uint32_t)a.spp + k, 0u, ((uint64_t)v.seed_hi << 32) | v.seed_lo);
                    const float xu = rng_next(&rng);
                    const float xv = rng_next(&rng);
                    // radiance_paths: normalised direction, box entry, or the sky
                    ori = v.ori;
                    dir = normalize(camera_dir(v.forward, v.right, v.up, row, col, v.width, v.height, xu, xv));
                    const float dis = ray_box_offset(ori, dir);
                    t_first = -1.0f;
                    if (dis < 0.0f) {
                        const vec3 sky = env_radiance(v.env, dir);
                        out[it] = make_float4(sky.x, sky.y, sky.z, -1.0f);
                    } else {
                        pos = ori + dir * dis;
                        csr = v3(1.0f, 1.0f, 1.0f);
                        res = v3(0.0f, 0.0f, 0.0f);
                        bounce = 0;
                        active = v.multi_scatter > 0;
                        if (!active) out[it] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                    }
                }
            }
        }
        if (queue_done && !__any(active)) break;  // uniform: every wave reaches it once its lanes are idle
        if (active) {
            const FallbackView& v = views[view];
            // several volumes: no brick occupancy (read from global memory it costs more than it saves)
            const CellsFetch F{v.cells, v.res, Occupancy{nullptr, 0u, 0u}};
            const vec3 light = normalize(v.light);
            vec3 np;
            bool done = false;
            if (!free_flight_batched<kFallbackBatch>(&rng, F, v.mj.inv_smax, v.mj.ratio, pos, dir,
                                                     ray_box_distance(pos, dir), kFallbackSteps, &np, &lk)) {
                res = res + env_radiance(v.env, dir) * csr;
                done = true;
            } else {
                rng_seek(&rng, rng.n - 2u);  // the two direction draws the free flight consumed
                const float e0 = rng_next(&rng);
                const float e1 = rng_next(&rng);
                const vec3 nd = hg_sample_dir(e0, e1, dir, v.g);
                if (bounce == 0) t_first = distance(ori, np);
                pos = np;
                csr = csr * v.scatter_rate;
                const float phase = hg_phase(dot(dir, light), v.g);
                const float tr = transmittance_batched<kFallbackBatch>(&rng, F, v.mj.inv_smax_tr, v.mj.ratio, pos,
                                                                       light, ray_box_distance(pos, light),
                                                                       kFallbackSteps, &lk);
                res = res + v.light_color * csr * (tr * phase);
                dir = nd;
                done = ++bounce >= v.multi_scatter;
            }
            if (done) {
                const vec3 r = res / 1.0f;  // radiance_paths' mean over its one sample
                out[it] = make_float4(r.x, r.y, r.z, t_first);
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
    if (threadIdx.x == 0 && red[0]) atomicAdd(&counters[3], (unsigned long long)red[0]);
}

// One thread per fallback pixel of the pass: its samples, in order, replace
// the MRPNN mean and histogram.
__global__ __launch_bounds__(256) void fallback_finish_kernel(ServeArgs a, const uint32_t* __restrict__ list,
                                                              uint32_t j0, uint32_t n_pix,
                                                              const float4* __restrict__ samples,
                                                              float* __restrict__ rgb, float* __restrict__ histo) {
    const uint32_t jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= n_pix) return;
    const size_t gi = list[j0 + jj];
    float tg[3] = {0.0f, 0.0f, 0.0f};
    float* hp = histo + (size_t)kHistoFloats * gi;
    for (int k = 0; k < a.fallback_spp; ++k) {
        const float4 r = samples[(size_t)jj * a.fallback_spp + k];
        histo_accumulate(tg, hp, vmax(v3(0, 0, 0), v3(r.x, r.y, r.z)), k);
    }
    rgb[3 * gi] = tg[0];
    rgb[3 * gi + 1] = tg[1];
    rgb[3 * gi + 2] = tg[2];
}

static vec3 V(const mrp_vec3& a) { return v3(a.x, a.y, a.z); }

static int ensure_serve_scratch(Volume* v, size_t bytes, void** out) {
    if (v->serve_scratch_bytes < bytes) {
        if (v->serve_scratch) {
            MRP_HIP(hipStreamSynchronize(v->stream));
            MRP_HIP(hipFree(v->serve_scratch));
        }
        v->serve_scratch = nullptr;
        v->serve_scratch_bytes = 0;
        if (hipMalloc(&v->serve_scratch, bytes) != hipSuccess) return fail(MRP_E_NOMEM, "serving buffers allocation failed");
        v->serve_scratch_bytes = bytes;
    }
    *out = v->serve_scratch;
    return MRP_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_serve_views(mrp_volume* const* volumes, const mrp_render_params* views, int n_views,
                               const mrp_serve_params* sp, float* rgb_dev, uint32_t* bgra_dev, uint8_t* flags_dev,
                               mrp_serve_stats* stats) {
    if (!volumes || !views || !sp || !rgb_dev) return fail(MRP_E_INVALID, "NULL argument");
    if (n_views <= 0 || n_views > 4096) return fail(MRP_E_INVALID, "n_views must be in [1, 4096]");
    if (sp->mrpnn_spp < 1 || sp->mrpnn_spp > 256) return fail(MRP_E_INVALID, "mrpnn_spp must be in [1, 256]");
    if (sp->fallback_spp < 1 || sp->fallback_spp > (1 << 16))
        return fail(MRP_E_INVALID, "fallback_spp must be in [1, 65536]");
    if (!(sp->var_threshold >= 0.0f)) return fail(MRP_E_INVALID, "var_threshold must be >= 0");
    if (sp->tone < 0 || sp->tone > 2) return fail(MRP_E_INVALID, "tone must be 0, 1 or 2");
    std::vector<Volume*> vols((size_t)n_views);
    for (int k = 0; k < n_views; ++k) {
        if (!volumes[k]) return fail(MRP_E_INVALID, "NULL volume");
        vols[k] = &volumes[k]->v;
        int rc = check_params(vols[k], &views[k]);
        if (rc) return rc;
        if (views[k].render_type != MRP_RENDER_MRPNN) return fail(MRP_E_INVALID, "views must be MRPNN renders");
    }
    Volume* host = vols[0];
    const int W = views[0].width, H = views[0].height;
    const size_t P = (size_t)W * H, n_pix = (size_t)n_views * P;
    if (n_pix >= ((size_t)1 << 32) || n_pix * sp->mrpnn_spp >= ((size_t)1 << 32) - 1)
        return fail(MRP_E_INVALID, "batch too large (more than 2^32 pixel samples)");

    // serving buffers in the host volume: view tables, MRPNN samples, histograms, fallback list
    const size_t b_tables = align256(mrpnn_view_table_bytes(n_views));
    const size_t b_fb = align256((size_t)n_views * sizeof(FallbackView));
    const size_t b_samples = align256(n_pix * (size_t)sp->mrpnn_spp * sizeof(float4));
    const size_t b_histo = align256(n_pix * kHistoFloats * sizeof(float));
    const size_t b_list = align256(n_pix * sizeof(uint32_t));
    // fallback samples of one pass: at most kFallbackPassItems (pixel, sample) items
    const size_t pass_items = std::min(n_pix * (size_t)sp->fallback_spp, kFallbackPassItems);
    const size_t b_fs = align256(pass_items * sizeof(float4));
    void* scratch = nullptr;
    int rc = ensure_serve_scratch(host, b_tables + b_fb + b_samples + b_histo + b_list + b_fs + 256, &scratch);
    if (rc) return rc;
    char* sb = (char*)scratch;
    void* tables = sb;
    FallbackView* fb_dev = (FallbackView*)(sb + b_tables);
    float4* samples = (float4*)(sb + b_tables + b_fb);
    float* histo = (float*)(sb + b_tables + b_fb + b_samples);
    uint32_t* list = (uint32_t*)(sb + b_tables + b_fb + b_samples + b_histo);
    float4* fsamples = (float4*)(sb + b_tables + b_fb + b_samples + b_histo + b_list);
    uint32_t* qc = (uint32_t*)(sb + b_tables + b_fb + b_samples + b_histo + b_list + b_fs);  // [count, cursor]

    for (hipEvent_t& e : host->serve_events)
        if (!e) MRP_HIP(hipEventCreate(&e));
    hipEvent_t* ev = host->serve_events;
    MRP_HIP(hipEventRecord(ev[0], host->stream));
    rc = clear_counters(host);
    if (rc) return rc;
    NNTimes tm;
    uint64_t n_active = 0;
    rc = mrpnn_views(vols.data(), views, n_views, sp->mrpnn_spp, tables, samples, &tm, &n_active);
    if (rc) return rc;

    std::vector<FallbackView> fb((size_t)n_views);
    for (int k = 0; k < n_views; ++k) {
        const Volume* v = vols[k];
        const mrp_render_params& p = views[k];
        FallbackView& f = fb[k];
        f.cells = v->cells;
        f.res = v->res;
        f.mj = make_majorant(v->max_density, p.alpha, p.tr_scale);
        f.g = p.g;
        f.ori = V(p.ori);
        f.forward = normalize(-f.ori);
        f.right = V(p.right);
        f.up = V(p.up);
        f.light = V(p.light_dir);
        f.light_color = V(p.light_color);
        f.scatter_rate = V(p.scatter_rate);
        f.env = env_map(v, V(p.sky));
        f.width = W;
        f.height = H;
        f.multi_scatter = p.multi_scatter;
        f.seed_lo = (uint32_t)p.seed;
        f.seed_hi = (uint32_t)(p.seed >> 32);
    }
    ServeArgs a;
    a.n_views = (uint32_t)n_views;
    a.P = (uint32_t)P;
    a.spp = sp->mrpnn_spp;
    a.fallback_spp = sp->fallback_spp;
    a.threshold = sp->var_threshold;
    MRP_HIP(hipMemcpyAsync(fb_dev, fb.data(), fb.size() * sizeof(FallbackView), hipMemcpyHostToDevice, host->stream));
    MRP_HIP(hipEventRecord(ev[1], host->stream));
    clear_u32_pair_kernel<<<1, 64, 0, host->stream>>>(qc);
    MRP_HIP(hipGetLastError());
    serve_accumulate_kernel<<<(unsigned)((n_pix + 255) / 256), 256, 0, host->stream>>>(a, samples, rgb_dev, histo,
                                                                                       flags_dev, list, qc);
    MRP_HIP(hipGetLastError());
    // the fallback count decides the passes: one read-back
    MRP_HIP(hipMemcpyAsync(host->host_words, qc, sizeof(uint32_t), hipMemcpyDeviceToHost, host->stream));
    MRP_HIP(hipStreamSynchronize(host->stream));
    const uint32_t n_flag = (uint32_t)host->host_words[0];
    MRP_HIP(hipEventRecord(ev[2], host->stream));
    const unsigned fb_blocks = (unsigned)std::max(1, host->num_cus * kFallbackBlocksPerCU);
    const uint32_t pass_pix = (uint32_t)std::max<size_t>(1, pass_items / (size_t)sp->fallback_spp);
    for (uint32_t j0 = 0; j0 < n_flag; j0 += pass_pix) {
        const uint32_t np = std::min(pass_pix, n_flag - j0);
        const uint32_t items = np * (uint32_t)sp->fallback_spp;
        clear_u32_pair_kernel<<<1, 64, 0, host->stream>>>(qc);  // the cursor (the count was read)
        MRP_HIP(hipGetLastError());
        fallback_kernel<<<fb_blocks, 256, 0, host->stream>>>(a, fb_dev, list, j0, items, qc + 1, fsamples,
                                                                    host->counters);
        MRP_HIP(hipGetLastError());
        fallback_finish_kernel<<<(np + 255) / 256, 256, 0, host->stream>>>(a, list, j0, np, fsamples, rgb_dev, histo);
        MRP_HIP(hipGetLastError());
    }
    MRP_HIP(hipEventRecord(ev[3], host->stream));
    if (bgra_dev) {
        rc = launch_denoise_images(host->stream, W, H, n_views, sp->denoise ? 1 : 0, sp->tone, sp->exposure, rgb_dev,
                                   histo, bgra_dev, std::max(sp->mrpnn_spp, sp->fallback_spp));
        if (rc) return rc;
    }
    MRP_HIP(hipEventRecord(ev[4], host->stream));
    // counters into pinned memory, then one synchronisation
    MRP_HIP(hipMemcpyAsync(host->host_words + 1, host->counters, 4 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, host->stream));
    MRP_HIP(hipStreamSynchronize(host->stream));
    if (stats) {
        *stats = mrp_serve_stats{};
        stats->mrpnn_lookups = host->host_words[1] + host->host_words[3];
        stats->fallback_lookups = host->host_words[4];
        stats->descriptors = n_active;
        stats->fallback_pixels = n_flag;
        MRP_HIP(hipEventElapsedTime(&stats->total_ms, ev[0], ev[4]));
        MRP_HIP(hipEventElapsedTime(&stats->mrpnn_ms, ev[0], ev[1]));
        MRP_HIP(hipEventElapsedTime(&stats->accumulate_ms, ev[1], ev[2]));
        MRP_HIP(hipEventElapsedTime(&stats->fallback_ms, ev[2], ev[3]));
        MRP_HIP(hipEventElapsedTime(&stats->denoise_ms, ev[3], ev[4]));
    }
    return MRP_OK;
}
