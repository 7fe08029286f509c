// mrp_tracer.hip - delta-tracking volume path tracer for gfx950.
//
// Reference algorithm: `CalculateRadiance` / `DeterminateNextVertex` / `Tr`
// (core/render.cuh:93-258) driven by the offline `RenderCamera<false>` kernel
// and `VolumeRender::Render` (core/volume.cu:147-179, 410-467).
//
// Structure (MI355X-first, not a translation of the reference's
// one-thread-per-pixel launch-per-sample loop):
//   * One persistent launch covers every (pixel, sample) path of a batch.
//     Workgroup b reads work queue b % 8, i.e. the queue of the XCD the
//     workgroup is dispatched to, and each queue owns one horizontal band of
//     the image (8x8-pixel tiles, all samples of a tile together), so rays that share an XCD's
//     L2 also share voxels.  A wave that drains its queue steals from the next.
//   * Lanes fetch items in wave-batched atomics (one atomicAdd per wave) and
//     regenerate a new path as soon as theirs ends, so lanes stay busy despite
//     path lengths that differ by orders of magnitude.
//   * Each path is a state machine (INIT, FF = free flight, TR1 = ratio
//     tracking, TR2 = delta tracking toward the light).  Every loop iteration
//     advances every live lane by exactly one tracking step, and all lanes
//     reach the single density-lookup site together, so the eight corner reads
//     issue as one wide memory instruction stream instead of three divergent ones.
//   * Random numbers: Philox4x32-10 per (pixel, sample); each lane's words
//     sit in a 16-word ring in LDS, topped up at one site per iteration, so
//     the cipher code is not replicated at every draw site.
//   * Results go to a per-(sample, pixel) float4 buffer; a second kernel folds
//     them into the running mean in sample order, which keeps the image
//     bit-identical to a sequential per-pixel loop (the CPU oracle).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mrp_internal.h"

namespace mrp {

// Path states: idle lane, pending path start, free flight, ratio tracking and
// delta tracking toward the light, pending real collision, pending end of the
// shadow ray, finished.
enum : int { ST_IDLE = 0, ST_INIT = 1, ST_FF = 2, ST_TR1 = 3, ST_TR2 = 4, ST_ACC = 5, ST_TRD = 6, ST_DONE = 7 };

struct TraceArgs {
    const float4* cells;
    Occupancy occ;  // brick occupancy of the cells: lookups in clear bricks read no grid line
    int res;
    float smax, alpha, g, tr_scale;
    vec3 ori, forward, right, up, light_dir, light_color, scatter_rate;
    EnvMap env;  // `SkyBox`: the environment map, or the constant sky
    Majorant mj;
    int width, height, multi_scatter;
    int checker, rows_up, wl;  // pixel selection (checker_col) and camera row convention; wl = logical width
    uint32_t seed_lo, seed_hi;
    int spp_begin, spp_count;
    int band_rows, tiles_x, tiles_per_band;
    uint32_t items_per_queue;
    uint32_t* queues;
    unsigned long long* counters;
    float4* out;
};

// Each lane's Philox words live in a ring of kRing words in LDS: word i of
// the lane at ring[(i % kRing) * 64] (the 64 lanes of a wave interleaved, so
// a wave's reads of one word each hit 64 different banks whatever word each
// lane reads).  `n` is the next draw index, `hi` the number of words
// generated so far.  An iteration takes at most 4 draws; a block of 4 is
// generated for every lane with room (hi - n <= kRing - 4) whenever some lane
// of the wave has fewer than 4 left.  With 8 words some lane needed a block
// in 97 % of the iterations; with 16 the wave runs a block in about 58 % of
// them (the lanes that top up together stay in step), and the ring costs no
// VGPRs.  The draws themselves do not depend on when a block is made.
constexpr uint32_t kRing = 16u;
struct Ring {
    uint32_t* w;  // this lane's word slot 0
    uint32_t n, hi;
};

__device__ __forceinline__ float ring_take(Ring& r) {
    const uint32_t x = r.w[(r.n & (kRing - 1u)) * 64u];
    r.n += 1u;
    return u01_(x);
}

__device__ __forceinline__ float fetch_density(const TraceArgs& a, vec3 p) {
    return trilinear_cells(a.cells, a.res, p, a.occ);
}

// Pending-event batching.  An event (path start, real collision, end of a
// shadow ray) runs a long divergent block (camera ray and box entry; HG
// direction sampling; box exit along the new direction).  With one event per
// ~25 steps per lane, some lane of 64 needs each block in almost every
// iteration, so executing it eagerly costs the whole wave several hundred
// instructions per lookup.  Instead a lane that hits an event parks in a
// pending state and the block runs when at least kEventBatch lanes wait for
// it, or when fewer than kMinStepping lanes are still stepping.
constexpr int kEventBatch = 16;
constexpr int kMinStepping = 40;

// kEnvMap: escaping paths read the environment map (a.env.px); without one
// they add the constant a.env.sky.  Both forms are held to 4 waves per SIMD
// (at most 128 VGPRs, no spills; left alone the compiler took 130 with the
// brick-occupancy test and 136 with the map, 3 waves).
template <bool kEnvMap>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void trace_kernel(TraceArgs a) {
    extern __shared__ uint32_t s_occ[];                 // occupancy_words(a.res) words
    __shared__ uint32_t s_ring[(256 / 64) * kRing * 64];  // one ring per lane (16 KB)
    a.occ = stage_occupancy(a.occ, a.res, s_occ);
    const uint32_t lane = __lane_id();
    int q = (int)(blockIdx.x & 7u);
    int qtries = 0;

    int state = ST_IDLE;
    int pix = 0, sidx = 0;
    Ring rng;
    rng.w = s_ring + (threadIdx.x >> 6) * (kRing * 64u) + lane;
    rng.n = rng.hi = 0u;
    vec3 pos = v3(0, 0, 0), rdir = v3(0, 0, 0), ndir = v3(0, 0, 0);
    vec3 csr = v3(1, 1, 1), res = v3(0, 0, 0);
    float t = 0.0f, dis = 0.0f, tfirst = -1.0f, tr = 1.0f, light_phase = 0.0f;
    int scatter_num = 0, loop_num = 0;
    uint32_t n_lookups = 0, n_paths = 0, n_scatters = 0;
    uint32_t path_iters = 0, max_iters = 0;  // loop iterations of the current / longest path

    const int P = a.width * a.height;
    // Wave-private slice [wq_cur, wq_end) of queue wq, claimed kChunk items at a
    // time so that the device-scope atomic (a round trip of microseconds across
    // XCDs) is paid once per kChunk paths rather than once per iteration.
    constexpr uint32_t kChunk = 128u;
    uint32_t wq_cur = 0u, wq_end = 0u;
    int wq = q;

    for (;;) {
        const bool stepping = (state == ST_FF) | (state == ST_TR1) | (state == ST_TR2);
        const int n_stepping = __popcll(__ballot(stepping));
        const bool flush = n_stepping < kMinStepping;

        // ------------------------------------------------ refill idle lanes
        const bool want = (state == ST_IDLE);
        const unsigned long long mask = __ballot(want);
        if (mask != 0ull && (flush || __popcll(mask) >= kEventBatch)) {
            if (wq_cur == wq_end && qtries < kNumQueues) {
                uint32_t base = 0u;
                if (lane == 0u) base = atomicAdd(&a.queues[q], kChunk);
                base = __shfl(base, 0);
                if (base >= a.items_per_queue) {  // this queue is exhausted: move on
                    q = (q + 1) & 7;
                    qtries += 1;
                } else {
                    wq = q;
                    wq_cur = base;
                    wq_end = min(base + kChunk, a.items_per_queue);
                }
            }
            if (wq_cur < wq_end) {
                const int rank = __popcll(mask & ((1ull << lane) - 1ull));
                const uint32_t take = min(wq_end - wq_cur, (uint32_t)__popcll(mask));
                const uint32_t base = wq_cur;
                const int qf = wq;
                wq_cur += take;
                if (want && (uint32_t)rank < take) {
                    const uint32_t k = base + (uint32_t)rank;
                    // tile-major: all samples of one 8x8 tile are consecutive items, so
                    // the paths in flight on an XCD start from a compact patch of image
                    const uint32_t per_tile = (uint32_t)a.spp_count * 64u;
                    const uint32_t tile = k / per_tile;
                    const uint32_t r = k - tile * per_tile;
                    const uint32_t s = r >> 6, w = r & 63u;
                    const int ty = (int)(tile / (uint32_t)a.tiles_x), tx = (int)(tile % (uint32_t)a.tiles_x);
                    const int col_l = tx * 8 + (int)(w & 7u);
                    const int row0 = qf * a.band_rows;
                    const int row = row0 + ty * 8 + (int)(w >> 3);
                    const int row_end = min(a.height, row0 + a.band_rows);
                    const int col = checker_col(row, col_l, a.checker);
                    if (col_l < a.wl && col < a.width && row < row_end) {
                        pix = row * a.width + col;
                        sidx = (int)s;
                        state = ST_INIT;
                        path_iters = 0;
                        rng.n = 0u;
                        rng.hi = 0u;
                    }
                }
            }
        }
        if (__ballot(state != ST_IDLE) == 0ull) {
            if (qtries >= kNumQueues && wq_cur == wq_end) break;
            continue;
        }

        // ------------------------------------------------ RNG top-up (one site)
        if (__ballot(state != ST_IDLE && rng.hi - rng.n < 4u) != 0ull && state != ST_IDLE &&
            rng.hi - rng.n <= kRing - 4u) {
            u32x4 c;
            c.v[0] = (uint32_t)pix;
            c.v[1] = (uint32_t)(a.spp_begin + sidx);
            c.v[2] = rng.hi >> 2;
            c.v[3] = 0u;
            const u32x4 o = philox4x32_10(c, a.seed_lo, a.seed_hi);
            uint32_t* d = rng.w + (rng.hi & (kRing - 1u)) * 64u;  // slots hi..hi+3 (hi is a multiple of 4)
            d[0] = o.v[0];
            d[64] = o.v[1];
            d[128] = o.v[2];
            d[192] = o.v[3];
            rng.hi += 4u;
        }
        if (state != ST_IDLE) path_iters++;

        // ------------------------------------------------ path starts
        {
            const unsigned long long m = __ballot(state == ST_INIT);
            if (m != 0ull && (flush || __popcll(m) >= kEventBatch) && state == ST_INIT) {
                // camera ray (core/volume.cu:158-163), CalculateRadiance entry (core/render.cuh:205-226)
                const float xu = ring_take(rng);
                const float xv = ring_take(rng);
                const int row = pix / a.width, col = pix - row * a.width;
                vec3 dir = camera_dir_rows(a.forward, a.right, a.up, row, col, a.width, a.height, xu, xv, a.rows_up);
                dir = normalize(dir);
                const float off = ray_box_offset(a.ori, dir);
                tfirst = -1.0f;
                if (off < 0.0f) {
                    res = kEnvMap ? env_radiance(a.env, dir) : a.env.sky;
                    state = ST_DONE;
                } else {
                    pos = a.ori + dir * off;
                    rdir = dir;
                    csr = v3(1, 1, 1);
                    res = v3(0, 0, 0);
                    scatter_num = 0;
                    if (a.multi_scatter <= 0) {
                        state = ST_DONE;
                    } else {
                        dis = ray_box_distance(pos, rdir);
                        t = 0.0f;
                        loop_num = 0;
                        state = ST_FF;
                    }
                }
            }
        }
        // ------------------------------------------------ real collisions (DeterminateNextVertex tail)
        {
            const unsigned long long m = __ballot(state == ST_ACC);
            if (m != 0ull && (flush || __popcll(m) >= kEventBatch) && state == ST_ACC) {
                const float e0 = ring_take(rng);
                const float e1 = ring_take(rng);
                ndir = hg_sample_dir(e0, e1, rdir, a.g);
                const vec3 npos = (rdir * t) + pos;
                n_scatters++;
                if (scatter_num == 0) tfirst = distance(a.ori, npos);
                pos = npos;
                csr = csr * a.scatter_rate;
                dis = ray_box_distance(pos, a.light_dir);
                light_phase = hg_phase(dot(rdir, a.light_dir), a.g);
                tr = 1.0f;
                t = 0.0f;
                loop_num = 0;
                state = ST_TR1;
            }
        }
        // ------------------------------------------------ shadow rays done (next-event estimate)
        {
            const unsigned long long m = __ballot(state == ST_TRD);
            if (m != 0ull && (flush || __popcll(m) >= kEventBatch) && state == ST_TRD) {
                res = res + a.light_color * csr * (tr * light_phase);
                rdir = ndir;
                scatter_num++;
                if (scatter_num < a.multi_scatter) {
                    dis = ray_box_distance(pos, rdir);
                    t = 0.0f;
                    loop_num = 0;
                    state = ST_FF;
                } else {
                    state = ST_DONE;
                }
            }
        }
        // ------------------------------------------------ finished paths
        if (state == ST_DONE) {
            a.out[(size_t)sidx * (size_t)P + (size_t)pix] = make_float4(res.x, res.y, res.z, tfirst);
            n_paths++;
            max_iters = max(max_iters, path_iters);
            state = ST_IDLE;
        }

        // ------------------------------------------------ one tracking step
        bool lookup = false;
        vec3 lp = v3(0, 0, 0);
        float rk2 = 0.0f;
        if ((state == ST_FF) | (state == ST_TR1) | (state == ST_TR2)) {
            // loop-condition checks of the reference while loops
            bool go = true;
            if (state == ST_FF) {
                const bool c = loop_num < 10000;
                loop_num++;
                if (!c) {
                    state = ST_ACC;  // loop cap: scatter at the current distance
                    go = false;
                }
            } else {
                if (state == ST_TR1) {
                    const bool c = loop_num < 10000;
                    loop_num++;
                    if (!(c && tr > 0.5f)) state = ST_TR2;
                }
                if (state == ST_TR2) {
                    const bool c = loop_num < 10000;
                    loop_num++;
                    if (!(c && tr > 0.001f)) {
                        state = ST_TRD;
                        go = false;
                    }
                }
            }
            if (go) {
                const float rk = ring_take(rng);
                t -= logf_(1.0f - rk) * (state == ST_FF ? a.mj.inv_smax : a.mj.inv_smax_tr);
                if (t > dis) {
                    if (state == ST_FF) {
                        res = res + (kEnvMap ? env_radiance(a.env, rdir) : a.env.sky) * csr;  // escaped: environment
                        state = ST_DONE;
                    } else {
                        if (state == ST_TR1) {
                            // the delta-tracking loop still runs once: its check, one
                            // distance draw, then the same overshoot breaks it
                            const bool c = loop_num < 10000;
                            loop_num++;
                            if (c && tr > 0.001f) (void)ring_take(rng);
                        }
                        state = ST_TRD;
                    }
                } else {
                    lookup = true;
                    lp = pos + ((state == ST_FF ? rdir : a.light_dir) * t);
                    if (state == ST_FF) rk2 = ring_take(rng);
                }
            }
        }

        // ------------------------------------------------ the one lookup site
        if (lookup) {
            const float d = fetch_density(a, lp);
            n_lookups++;
            const float ratio = d * a.mj.ratio;
            if (state == ST_FF) {
                if (ratio > rk2) state = ST_ACC;
                else if (d < 0.0f) t -= d;
            } else if (state == ST_TR1) {
                tr *= 1.0f - fmax_(0.0f, ratio);
                if (d < 0.0f) t -= d;
            } else {
                const float rk3 = ring_take(rng);
                if (rk3 < fmax_(0.0f, ratio)) {
                    tr = 0.0f;
                    state = ST_TRD;
                } else if (d < 0.0f) {
                    t -= d;
                }
            }
        }
        if (state == ST_DONE) {
            a.out[(size_t)sidx * (size_t)P + (size_t)pix] = make_float4(res.x, res.y, res.z, tfirst);
            n_paths++;
            max_iters = max(max_iters, path_iters);
            state = ST_IDLE;
        }
    }

    if (n_lookups) atomicAdd(&a.counters[0], (unsigned long long)n_lookups);
    if (n_paths) atomicAdd(&a.counters[1], (unsigned long long)n_paths);
    if (n_scatters) atomicAdd(&a.counters[2], (unsigned long long)n_scatters);
    if (max_iters) atomicMax(&a.counters[3], (unsigned long long)max_iters);
}

// Fold per-sample results into the running mean in sample order
// (`target = lerp(target, max(0,r), 1/(1+s))`, core/volume.cu:173-178) and
// optionally into FP64 sums.
__global__ __launch_bounds__(256) void accumulate_kernel(const float4* __restrict__ samples, int P, int spp_begin,
                                                         int spp_count, float* __restrict__ target,
                                                         double* __restrict__ sum64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    vec3 tg = v3(target[3 * i], target[3 * i + 1], target[3 * i + 2]);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < spp_count; ++k) {
        const float4 r = samples[(size_t)k * (size_t)P + (size_t)i];
        const vec3 c = vmax(v3(0, 0, 0), v3(r.x, r.y, r.z));
        const float rate = 1.0f / (float)(1 + spp_begin + k);
        tg = lerp(tg, c, rate);
        s0 += (double)c.x;
        s1 += (double)c.y;
        s2 += (double)c.z;
    }
    target[3 * i] = tg.x;
    target[3 * i + 1] = tg.y;
    target[3 * i + 2] = tg.z;
    if (sum64) {
        sum64[3 * i] += s0;
        sum64[3 * i + 1] += s1;
        sum64[3 * i + 2] += s2;
    }
}

__global__ void clear_u32_kernel(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void clear_u64_kernel(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

static vec3 V(const mrp_vec3& a) { return v3(a.x, a.y, a.z); }

int launch_accumulate(Volume* v, const void* samples, size_t P, int spp_begin, int n, float* target, double* sum64) {
    accumulate_kernel<<<(unsigned)((P + 255) / 256), 256, 0, v->stream>>>((const float4*)samples, (int)P, spp_begin, n,
                                                                       target, sum64);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

int check_params(const Volume* v, const mrp_render_params* p) {
    if (!p) return fail(MRP_E_INVALID, "params is NULL");
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() must run before rendering");
    if (p->width <= 0 || p->height <= 0 || (long long)p->width * p->height > (1ll << 28))
        return fail(MRP_E_INVALID, "image size out of range");
    if (p->spp_begin < 0 || p->spp_count <= 0 || p->spp_begin > (1 << 24) || p->spp_count > (1 << 24))
        return fail(MRP_E_INVALID, "sample range out of range");
    if (!(p->alpha > 0.0f)) return fail(MRP_E_INVALID, "alpha must be > 0");
    if (p->multi_scatter < 0) return fail(MRP_E_INVALID, "multi_scatter must be >= 0");
    if (p->render_type != MRP_RENDER_PT && p->render_type != MRP_RENDER_MRPNN)
        return fail(MRP_E_INVALID, "render_type must be MRP_RENDER_PT or MRP_RENDER_MRPNN");
    return MRP_OK;
}

static void fill_trace_args(Volume* v, const mrp_render_params* p, PixelMap map, TraceArgs* a) {
    const int W = p->width, H = p->height;
    const int wl = map.checker < 0 ? W : (W + 1) / 2;
    const int band_rows = (H + kNumQueues - 1) / kNumQueues;
    a->cells = v->cells;
    a->occ = Occupancy{v->occ, occupancy_shift(v->res), occupancy_n(v->res)};
    a->res = v->res;
    a->alpha = p->alpha;
    a->smax = v->max_density * p->alpha;
    a->mj = make_majorant(v->max_density, p->alpha, p->tr_scale);
    a->g = p->g;
    a->tr_scale = p->tr_scale;
    a->ori = V(p->ori);
    a->forward = normalize(-a->ori);
    a->right = V(p->right);
    a->up = V(p->up);
    a->light_dir = normalize(V(p->light_dir));
    a->light_color = V(p->light_color);
    a->env = env_map(v, V(p->sky));
    a->scatter_rate = V(p->scatter_rate);
    a->width = W;
    a->height = H;
    a->multi_scatter = p->multi_scatter;
    a->checker = map.checker;
    a->rows_up = map.rows_up;
    a->wl = wl;
    a->seed_lo = (uint32_t)p->seed;
    a->seed_hi = (uint32_t)(p->seed >> 32);
    a->band_rows = band_rows;
    a->tiles_x = (wl + 7) / 8;
    a->tiles_per_band = a->tiles_x * ((band_rows + 7) / 8);
    a->queues = v->queues;
    a->counters = v->counters;
    a->out = (float4*)v->samples;
}

static uint64_t trace_items_per_sample(const TraceArgs& a) { return (uint64_t)a.tiles_per_band * 64u; }

template <bool kEnvMap>
static int trace_grid(Volume* v, int* blocks) {
    int occ = 0;
    MRP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, trace_kernel<kEnvMap>, 256,
                                                         occupancy_words(v->res) * sizeof(uint32_t)));
    occ = std::max(1, std::min(occ, 8));
    *blocks = std::max(kNumQueues, v->num_cus * occ);
    return MRP_OK;
}

// One persistent tracer launch over samples [s_begin, s_begin + n) (absolute
// RNG sample indices) of the pixels that `map` selects; path results go to
// v->samples[(s - s_begin) * P + pixel].  The caller sized v->samples.
int trace_samples(Volume* v, const mrp_render_params* p, PixelMap map, int s_begin, int n) {
    TraceArgs a;
    fill_trace_args(v, p, map, &a);
    a.spp_begin = s_begin;
    a.spp_count = n;
    a.items_per_queue = (uint32_t)(trace_items_per_sample(a) * (uint64_t)n);
    clear_u32_kernel<<<1, 64, 0, v->stream>>>(v->queues, kNumQueues);
    int blocks = 0, rc;
    if (a.env.px) {
        rc = trace_grid<true>(v, &blocks);
        if (rc) return rc;
        trace_kernel<true><<<blocks, 256, occupancy_words(a.res) * sizeof(uint32_t), v->stream>>>(a);
    } else {
        rc = trace_grid<false>(v, &blocks);
        if (rc) return rc;
        trace_kernel<false><<<blocks, 256, occupancy_words(a.res) * sizeof(uint32_t), v->stream>>>(a);
    }
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

int clear_counters(Volume* v) {
    clear_u64_kernel<<<1, 64, 0, v->stream>>>(v->counters, 4);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

int read_counters(Volume* v, unsigned long long out[4]) {
    MRP_HIP(hipMemcpyAsync(out, v->counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, v->stream));
    MRP_HIP(hipStreamSynchronize(v->stream));
    return MRP_OK;
}

// Trace samples [spp_begin, spp_begin+spp_count) into target/sum64 (device).
static int trace_dev(Volume* v, const mrp_render_params* p, float* target, double* sum64, mrp_trace_stats* st) {
    int rc = check_params(v, p);
    if (rc) return rc;
    if (p->render_type == MRP_RENDER_MRPNN) return mrpnn_dev(v, p, target, sum64, st);
    const int W = p->width, H = p->height;
    const size_t P = (size_t)W * H;
    const size_t per_sample_bytes = P * sizeof(float4);
    const PixelMap map;
    TraceArgs probe;
    fill_trace_args(v, p, map, &probe);
    const uint64_t per_sample_items = trace_items_per_sample(probe);
    int batch = (int)std::max<size_t>(1, std::min<size_t>(v->sample_budget / per_sample_bytes, (size_t)p->spp_count));
    batch = (int)std::min<uint64_t>((uint64_t)batch, (uint64_t)(((1ull << 32) - (1ull << 26)) / per_sample_items));
    if (batch < 1) return fail(MRP_E_INVALID, "image too large for one work queue");
    rc = ensure_samples(v, (size_t)batch * per_sample_bytes);
    if (rc) return rc;
    int blocks = 0;
    rc = probe.env.px ? trace_grid<true>(v, &blocks) : trace_grid<false>(v, &blocks);
    if (rc) return rc;

    hipEvent_t e0, e1, e2, e3;
    MRP_HIP(hipEventCreate(&e0));
    MRP_HIP(hipEventCreate(&e1));
    MRP_HIP(hipEventCreate(&e2));
    MRP_HIP(hipEventCreate(&e3));
    double kernel_ms = 0.0, reduce_ms = 0.0;
    int launches = 0;
    MRP_HIP(hipEventRecord(e0, v->stream));
    rc = clear_counters(v);
    if (rc) return rc;
    for (int s0 = 0; s0 < p->spp_count; s0 += batch) {
        const int n = std::min(batch, p->spp_count - s0);
        const int sb = p->spp_begin + s0;
        MRP_HIP(hipEventRecord(e1, v->stream));
        rc = trace_samples(v, p, map, sb, n);
        if (rc) return rc;
        MRP_HIP(hipEventRecord(e2, v->stream));
        accumulate_kernel<<<(unsigned)((P + 255) / 256), 256, 0, v->stream>>>((const float4*)v->samples, (int)P, sb, n,
                                                                           target, sum64);
        MRP_HIP(hipGetLastError());
        MRP_HIP(hipEventRecord(e3, v->stream));
        MRP_HIP(hipEventSynchronize(e3));
        float ms = 0.0f;
        MRP_HIP(hipEventElapsedTime(&ms, e1, e2));
        kernel_ms += ms;
        MRP_HIP(hipEventElapsedTime(&ms, e2, e3));
        reduce_ms += ms;
        launches++;
    }
    float total = 0.0f;
    MRP_HIP(hipEventElapsedTime(&total, e0, e3));
    unsigned long long cnt[4] = {0, 0, 0, 0};
    rc = read_counters(v, cnt);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    hipEventDestroy(e3);
    if (rc) return rc;
    if (st) {
        st->lookups = cnt[0];
        st->paths = cnt[1];
        st->scatters = cnt[2];
        st->max_path_steps = cnt[3];
        st->kernel_ms = kernel_ms;
        st->reduce_ms = reduce_ms;
        st->total_ms = total;
        st->launches = launches;
        st->grid_blocks = blocks;
        st->first_vertex_ms = st->shadow_ms = st->compact_ms = st->gather_ms = 0.0;
        st->shadow_lookups = 0;
    }
    return MRP_OK;
}

}  // namespace mrp

using namespace mrp;

extern "C" int mrp_render_offline_dev(mrp_volume* vol, const mrp_render_params* p, float* target_dev,
                                      double* sum64_dev, mrp_trace_stats* stats) {
    if (!vol || !target_dev) return fail(MRP_E_INVALID, "volume/target is NULL");
    MRP_HIP(hipSetDevice(vol->v.device));
    return trace_dev(&vol->v, p, target_dev, sum64_dev, stats);
}

extern "C" int mrp_render_offline(mrp_volume* vol, const mrp_render_params* p, float* rgb_host,
                                  mrp_trace_stats* stats) {
    if (!vol || !rgb_host || !p) return fail(MRP_E_INVALID, "volume/params/output is NULL");
    if (p->width <= 0 || p->height <= 0) return fail(MRP_E_INVALID, "image size out of range");
    Volume* v = &vol->v;
    MRP_HIP(hipSetDevice(v->device));
    const size_t bytes = (size_t)p->width * p->height * 3 * sizeof(float);
    float* target = nullptr;
    MRP_HIP(hipMalloc(&target, bytes));
    int rc = MRP_OK;
    if (p->spp_begin > 0) {
        rc = hipMemcpy(target, rgb_host, bytes, hipMemcpyHostToDevice) == hipSuccess ? MRP_OK
                                                                                     : fail(MRP_E_HIP, "upload target");
    } else {
        rc = hipMemsetAsync(target, 0, bytes, v->stream) == hipSuccess ? MRP_OK : fail(MRP_E_HIP, "clear target");
    }
    if (rc == MRP_OK) rc = trace_dev(v, p, target, nullptr, stats);
    if (rc == MRP_OK && hipMemcpy(rgb_host, target, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(MRP_E_HIP, "download target");
    hipFree(target);
    return rc;
}
