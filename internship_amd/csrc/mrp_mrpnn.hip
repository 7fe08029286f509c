// mrp_mrpnn.hip - MRPNN inference path on gfx950 (FP32 reference-parity path).
//
// Reference: `NNPredict<MRPNN>` (core/render.cuh:469-523) and
// `RadiancePredict` (core/radiancePredict.cu:307-486), driven by the offline
// camera kernel (core/volume.cu:147-179).  The reference does everything for
// one pixel in one thread (about 7 KB of per-thread arrays).  Here the work is
// split into stages with their own launch shapes:
//
//   1. first_vertex_kernel : one thread per (sample, pixel) - first free flight
//      (g = 0), random view/light frames, HG to the light; compact_kernel lists
//      the vertices that scatter, and shadow_kernel traces their shadow rays
//      (single scatter Tr * HG) one thread per listed vertex, so the long
//      ratio-tracking loops run in full waves;
//   2. gather_kernel       : one thread per (point, tap) - 192 taps x
//      {density mip, transmittance mip, HG table} -> 576 features + gamma;
//   3. mlp (FP32 here, BF16 / FP8 MFMA in mrp_mlp_mfma.hip) - the network;
//   4. combine_kernel      : (predict + single) * light * albedo, or sky.
//
// The per-(sample, pixel) results then go through the same in-order running
// mean as the path tracer (launch_accumulate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "mrp_hglut.h"
#include "mrp_internal.h"
#include "mrp_mlp.h"
#include "mrp_mrpnn.h"
#include "mrp_track.h"
#include "mrp_weights_gen.h"

namespace mrp {

__constant__ float c_tap_offset[kTaps];
__constant__ int c_tap_mip[kTaps];
__constant__ float c_tap_v[kTaps];
__constant__ float c_stencil[kShellTaps * 3];
__constant__ float c_tap_inv_len[kShellTaps];
__constant__ float c_tap_u_view[kShellTaps];

struct NNArgs {
    const float4* cells;
    const float4* quads;  // the same grid in the quad layout (QuadsFetch), read by the first free flight
    QuadGeom qg;
    Occupancy occ;        // brick occupancy of the cells: lookups in clear bricks read no grid line
    int res;
    float smax, alpha, g, tr_scale;
    Majorant mj;
    vec3 ori, forward, right, up, light, lx, light_color, scatter_rate;
    EnvMap env;  // `SkyBox` of rays whose first free flight escapes
    int width, height;
    int checker, rows_up, wl;  // pixel selection (checker_col), camera row convention, logical width
    const float* ray_ori;      // a ray list instead of camera pixels (n x 3 each), or nullptr
    const float* ray_dir;
    uint32_t seed_lo, seed_hi;
    int spp_begin;       // absolute sample index of item 0's sample
    uint32_t item0;      // first (sample-major) item of this chunk within the batch
    uint32_t n_items;    // items in this chunk
    unsigned long long* counters;
    // several images in one launch: item k of the chunk belongs to view
    // (item0 + k) / view_items, whose record (camera, volume, light, seed) is
    // views[view]; nullptr for one image (this record)
    const NNArgs* views;
    uint32_t view_items;
    uint32_t occ_words;  // views: the largest view's occupancy words (the first vertex's LDS)
};

// ---------------------------------------------------------------- stage 1
// The first free flight and the shadow ray use the batched trackers of
// mrp_track.h: kFVBatch / kShadowBatch lookups in flight per thread, with the
// draws and results of the one-lookup-at-a-time loops.  The first vertex runs
// 2M threads, so occupancy times lookups in flight sets its memory
// parallelism.  With the brick occupancy in LDS most of its lookups in the
// cloud's hull hit L2, and more waves beat deeper batches: at 1080p 3 (93
// VGPRs, 5 waves per SIMD) took 0.36 ms against 4: 0.37, 6: 0.41, 8 (153
// VGPRs, 3 waves): 0.42; 12 put the batch in scratch (1.29 ms).  The shadow
// rays run about 256K threads, 4 waves per SIMD, so depth pays there: 6 (148 VGPRs) took 0.23 ms against 2: 0.26, 4: 0.25, 8: 0.24.
// The multi-view forms (serving: 64 views of 16 volumes, no brick occupancy)
// measured their batch at 4 and 4 (MRPNN stages 73.3 ms per 720p batch
// against 74.0 with 6 and 4 and 74.9 with 3 and 6 or 8 and 4).
// Shadow rays are taken in kShadowBuckets buckets of their expected length
// (shadow_bucket); the compaction counts them, the first vertex clears the counts.
constexpr int kShadowBuckets = 64;
constexpr int kFVBatch = 3;
constexpr int kShadowBatch = 6;
constexpr int kFVBatchViews = 4;
constexpr int kShadowBatchViews = 4;
constexpr int kMaxTrackSteps = 10000;

// pos_act = (pos, active), fx = (X, HG phase to the light; shadow_kernel
// multiplies in the transmittance) or, for a vertex that escaped, (the
// environment radiance along the ray, 0), fy = (Y, output slot as bits),
// fz = (Z, dist), rng_pos = the stream's draw count where the shadow ray
// starts.
// Items run over the logical pixels of `a`'s pixel map; the output slot
// (s - spp_begin) * P + pixel is 0xffffffff for a logical pixel past the image
// edge (odd widths under the checkerboard).
// count_out: the compaction's counter, cleared here (the compaction runs next
// on the same stream).
//
// first_vertex_item: one (sample, pixel) item; `item` counts from the first
// item of `a`'s image, `slot_base` is added to the output slot.
template <int kBatch>
__device__ __forceinline__ void first_vertex_item(const NNArgs& a, const Occupancy& occ, uint32_t i, uint32_t item,
                                                  uint32_t slot_base, float4* __restrict__ pos_act,
                                                  float4* __restrict__ fx, float4* __restrict__ fy,
                                                  float4* __restrict__ fz, uint32_t* __restrict__ rng_pos,
                                                  uint32_t* ring, uint32_t* lk) {
    const uint32_t P = (uint32_t)(a.width * a.height);
    const uint32_t Pl = (uint32_t)(a.wl * a.height);
    const uint32_t s = item / Pl, l = item - s * Pl;
    int row, col_l;
    if (a.ray_ori) {  // a ray list: item l is ray l
        row = 0;
        col_l = (int)l;
    } else {
        block_order(l, (uint32_t)a.wl, (uint32_t)a.height, &row, &col_l);  // waves cover 8 x 8 pixel blocks
    }
    const int col = checker_col(row, col_l, a.checker);
    if (col >= a.width) {
        pos_act[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        fx[i] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        fy[i] = make_float4(0.0f, 1.0f, 0.0f, __uint_as_float(0xffffffffu));
        fz[i] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
    } else {
        const uint32_t pix = (uint32_t)row * (uint32_t)a.width + (uint32_t)col;
        RingRng<64> rng =
            ring_make<64>(ring, pix, (uint32_t)a.spp_begin + s, 0u, ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        vec3 ori = a.ori, dir;
        if (a.ray_ori) {
            ori = v3(a.ray_ori[3 * (size_t)pix], a.ray_ori[3 * (size_t)pix + 1], a.ray_ori[3 * (size_t)pix + 2]);
            dir = v3(a.ray_dir[3 * (size_t)pix], a.ray_dir[3 * (size_t)pix + 1], a.ray_dir[3 * (size_t)pix + 2]);
        } else {
            const float xu = rng_next(&rng);
            const float xv = rng_next(&rng);
            dir = camera_dir_rows(a.forward, a.right, a.up, row, col, a.width, a.height, xu, xv, a.rows_up);
        }
        dir = normalize(dir);  // NNPredict
        const vec3 gdir = normalize(dir);  // GetSamplePoint
        vec3 pos = v3(0, 0, 0);
        int active = 0;
        const float off = ray_box_offset(ori, gdir);
        if (off >= 0.0f) {
            const vec3 sp = ori + gdir * off;
            vec3 np;
            const QuadsFetch F{a.quads, a.res, a.qg, occ};
            if (free_flight_batched<kBatch>(&rng, F, a.mj.inv_smax, a.mj.ratio, sp, gdir,
                                              ray_box_distance(sp, gdir), kMaxTrackSteps, &np, lk)) {
                pos = np;
                active = 1;
            }
        }
        // A vertex whose free flight escaped is sky: its frames and shadow ray
        // (the rest of NNPredict's draws) cannot change any output, so they are skipped.
        // The shadow ray of a scattering vertex runs in shadow_kernel over the
        // compacted vertices, from the draw count stored here.
        vec3 X = v3(1, 0, 0), Y = v3(0, 1, 0), Z = v3(0, 0, 1);
        float phase = 0.0f;
        if (active) {
            vec3 LX, LY, LZ;
            frame_from_normal(&rng, dir, &X, &Y, &Z);
            frame_from_normal(&rng, a.light, &LX, &LY, &LZ);
            phase = hg_phase(dot(X, LX), a.g);
            rng_pos[i] = rng.n;
        }
        const float dist = active ? distance(ori, pos) : -1.0f;
        pos_act[i] = make_float4(pos.x, pos.y, pos.z, active ? 1.0f : 0.0f);
        if (!active) X = env_radiance(a.env, dir);  // the sky this pixel sees (NNPredict's SkyBox(dir))
        fx[i] = make_float4(X.x, X.y, X.z, phase);
        fy[i] = make_float4(Y.x, Y.y, Y.z, __uint_as_float(slot_base + s * P + pix));
        fz[i] = make_float4(Z.x, Z.y, Z.z, dist);
    }
}

// kViews: the items run over the images of a.views (view_items items each,
// view-major); every view's count is a multiple of 64, so a wave's items
// belong to one view and its record is read with scalar loads.
template <bool kViews>
__global__ __launch_bounds__(256) void first_vertex_kernel(NNArgs a, float4* __restrict__ pos_act,
                                                           float4* __restrict__ fx, float4* __restrict__ fy,
                                                           float4* __restrict__ fz, uint32_t* __restrict__ rng_pos,
                                                           uint32_t* __restrict__ count_out,
                                                           uint32_t* __restrict__ buckets) {
    __shared__ unsigned red[256];
    __shared__ uint32_t s_ring[256 * kRingWords];  // the threads' Philox rings (16 KB)
    extern __shared__ uint32_t s_occ[];             // occupancy_words(a.res) words (one image), a.occ_words (views)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) *count_out = 0u;
    if (i < 2u * kShadowBuckets) buckets[i] = 0u;  // the compaction counts, the shadow stage's offsets
    uint32_t lk = 0u;
    if constexpr (kViews) {
        // several volumes: a workgroup whose items all belong to one view (all
        // but those that straddle two views) stages that view's brick
        // occupancy; the others read every cell (from global memory the bits
        // cost more than they save)
        const uint32_t first = a.item0 + blockIdx.x * blockDim.x;
        const uint32_t last = a.item0 + min((blockIdx.x + 1u) * blockDim.x, a.n_items) - 1u;
        const uint32_t v0 = first / a.view_items;
        const Occupancy occ = v0 == last / a.view_items  // uniform
                                  ? stage_occupancy(a.views[v0].occ, a.views[v0].res, s_occ)
                                  : Occupancy{nullptr, 0u, 0u};
        if (i < a.n_items) {
            const uint32_t item = a.item0 + i;
            const uint32_t view = __builtin_amdgcn_readfirstlane(item / a.view_items);
            const uint32_t base = view * a.view_items;
            first_vertex_item<kFVBatchViews>(a.views[view], occ, i, item - base, base, pos_act, fx, fy, fz, rng_pos,
                                             lds_ring(s_ring), &lk);
        }
    } else {
        const Occupancy occ = stage_occupancy(a.occ, a.res, s_occ);
        if (i < a.n_items)
            first_vertex_item<kFVBatch>(a, occ, i, a.item0 + i, 0u, pos_act, fx, fy, fz, rng_pos, lds_ring(s_ring), &lk);
    }
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&a.counters[0], (unsigned long long)red[0]);
}

// Shadow rays of the compacted scattering vertices (the `Tr` of NNPredict's
// single-scattering term, core/render.cuh:469-523), one thread per listed
// vertex: the stream resumes where first_vertex_kernel stopped, so every draw
// matches the one-thread-per-pixel order.  single[i] = Tr * phase (phase in
// fx.w).  Its lookups are counted in counters[2].
// The walk runs in rounds of kShadowBatch lookups whose speculation is shared by
// ratio and delta tracking, so lanes in different phases wait for memory
// together.  Measured alternatives, all slower: persistent lanes that claim
// the next vertex when done (0.78 against 0.27 ms at 1080p), 2-8 lanes per
// ray sharing one walk (0.35-0.64 ms), this kernel on a second stream
// beside the gather (both slowed by about what the overlap saved), and 8
// lookups in flight per lane (170 VGPRs, half the waves: 0.29 ms).
template <bool kViews>
__global__ __launch_bounds__(256) void shadow_kernel(NNArgs a0, const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ n_act_dev,
                                                     const float4* __restrict__ pos_act, const float4* __restrict__ fy,
                                                     const uint32_t* __restrict__ rng_pos,
                                                     const float4* __restrict__ fx, float* __restrict__ single) {
    __shared__ unsigned red[256];
    __shared__ uint32_t s_ring[256 * kRingWords];  // the threads' Philox rings (16 KB)
    extern __shared__ uint32_t s_occ[];             // occupancy_words(a0.res) words (one image), none (views)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_act = *n_act_dev;  // the compaction's count (the grid covers the whole chunk)
    if (blockIdx.x * blockDim.x >= n_act) return;  // uniform over the workgroup: no barrier is split
    const Occupancy occ = kViews ? Occupancy{nullptr, 0u, 0u} : stage_occupancy(a0.occ, a0.res, s_occ);
    uint32_t lk = 0u;
    if (k < n_act) {
        const uint32_t i = list[k];
        uint32_t slot = __float_as_uint(fy[i].w);
        // the vertex's view (lanes of a wave may belong to different views)
        const NNArgs* ap = &a0;
        if constexpr (kViews) {
            const uint32_t view = slot / a0.view_items;
            slot -= view * a0.view_items;
            ap = a0.views + view;
        }
        const NNArgs& a = *ap;
        const uint32_t P = (uint32_t)(a.width * a.height);
        const uint32_t s = slot / P, pix = slot - s * P;
        RingRng<64> rng = ring_make<64>(lds_ring(s_ring), pix, (uint32_t)a.spp_begin + s, 0u,
                                        ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        rng_seek(&rng, rng_pos[i]);
        const float4 pa = pos_act[i];
        const vec3 pos = v3(pa.x, pa.y, pa.z);
        const vec3 LX = normalize(a.light);  // the first axis frame_from_normal returns for the light
        const CellsFetch F{a.cells, a.res, occ};
        const float tr = transmittance_batched<kViews ? kShadowBatchViews : kShadowBatch>(
            &rng, F, a.mj.inv_smax_tr, a.mj.ratio, pos, LX, ray_box_distance(pos, LX), kMaxTrackSteps, &lk);
        single[i] = tr * fx[i].w;
    }
    red[threadIdx.x] = lk;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(&a0.counters[2], (unsigned long long)red[0]);
}

// Shadow-ray order: the listed vertices in buckets of their expected
// ratio-tracking steps, the exit distance toward the light times the
// transmittance majorant (log scale), longest first, so a wave's rays are of
// similar length and the longest start first (0.281 -> 0.261 ms at 1080p,
// the sort included).  Counting sort in two passes; the order within a
// bucket depends on scheduling, no result does.
template <bool kViews>
__device__ __forceinline__ int shadow_bucket(const NNArgs& a0, const float4* __restrict__ pos_act,
                                             const float4* __restrict__ fy, uint32_t i) {
    vec3 L = a0.light;
    float inv_maj = a0.mj.inv_smax_tr;
    if constexpr (kViews) {
        const NNArgs& av = a0.views[__float_as_uint(fy[i].w) / a0.view_items];
        L = av.light;
        inv_maj = av.mj.inv_smax_tr;
    }
    const float4 pa = pos_act[i];
    const float steps = ray_box_distance(v3(pa.x, pa.y, pa.z), normalize(L)) / inv_maj;
    const int b = (int)(__log2f(1.0f + steps) * 5.0f);  // 2^(64/5) steps and more share the last bucket
    return kShadowBuckets - 1 - (b < 0 ? 0 : (b > kShadowBuckets - 1 ? kShadowBuckets - 1 : b));
}


template <bool kViews>
__global__ __launch_bounds__(256) void shadow_scatter_kernel(NNArgs a0, const uint32_t* __restrict__ list,
                                                             const uint32_t* __restrict__ n_act_dev,
                                                             const float4* __restrict__ pos_act,
                                                             const float4* __restrict__ fy,
                                                             const uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ cursors,
                                                             uint32_t* __restrict__ order) {
    __shared__ uint32_t hist[kShadowBuckets];
    __shared__ uint32_t base[kShadowBuckets];
    const uint32_t n_act = *n_act_dev;
    if (blockIdx.x * blockDim.x >= n_act) return;  // uniform over the workgroup
    if (threadIdx.x < kShadowBuckets) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    int b = 0;
    uint32_t rank = 0u, i = 0u;
    if (k < n_act) {
        i = list[k];
        b = shadow_bucket<kViews>(a0, pos_act, fy, i);
        rank = atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kShadowBuckets) {
        uint32_t start = 0u;  // the bucket's first slot: the counts of the buckets before it
        for (int u = 0; u < (int)threadIdx.x; ++u) start += counts[u];
        base[threadIdx.x] = hist[threadIdx.x] ? start + atomicAdd(&cursors[threadIdx.x], hist[threadIdx.x]) : 0u;
    }
    __syncthreads();
    if (k < n_act) order[base[b] + rank] = i;
}

// ---------------------------------------------------------------- stage 2
struct GatherArgs {
    const float4* mips[kNumMips];    // border cells of the density mips
    const float4* trs[kNumTrMips];   // clamp cells of the transmittance mips
    const float4* lut;               // bilinear clamp cells of the HG table
    const uint4* mips16;             // FP16 cells of every density mip (tiled gather)
    const uint4* trs16;              // FP16 cells of every transmittance mip (tiled gather)
    vec3 light, lx;
    vec3 id, tf;  // inv3(lx) and 0.4999 * inv3(lx), RayBoxOffset's per-direction terms (made on the host: IEEE
                  // division and multiplication round alike there)
    float alpha, g;
    uint32_t n_items;
    const uint32_t* n_dev;  // if set, the descriptor count is read here (n_items bounds the grid)
    // several images in one launch: a descriptor's view is its output slot /
    // view_items and its tables are views[view]; nullptr for one image
    const GatherArgs* views;
    uint32_t view_items;
};

// Segments of the 12 squeeze-excitation blocks (`SE`/`SERES` call sites,
// core/radiancePredict.cu:396-434): tap range [from, from+size) per block, in
// pool order (block b writes means at 3b+k and maxes at 36+3b+k, k = stream).
__constant__ int c_seg_from[12] = {0, 8, 16, 32, 48, 64, 96, 128, 160, 168, 176, 184};
__constant__ int c_seg_size[12] = {8, 8, 16, 16, 16, 32, 32, 32, 8, 8, 8, 8};

// One thread per tap: the descriptor [Xd | Xt | Xh] (576 features),
// gamma[item] = acos(X . LX), and pool[item][72] = the per-segment means
// (0..35) and maxima (36..71) that the network's squeeze-excitation blocks
// start from (`AVGMAX`, core/radiancePredict.cu:21-33: ascending sum then
// /size; max from -10000).  Descriptors go out as FP32 rows (feat[item][576]),
// for the FP32 network and the callers of mrp_mrpnn_features; the MFMA
// renderers' gather is gather_tiled_kernel below.
// One workgroup (192 threads, one tap each) loops over kGatherItems points, so
// the dispatcher launches 1/kGatherItems as many workgroups; the LDS row is
// double-buffered, one barrier per point.  At 1080p (256k descriptors) 16
// took 0.59 ms, 8 took 0.61 ms and 32 took 0.62 ms.
constexpr int kGatherItems = 16;
static_assert(kGatherItems % 8 == 0, "the tiled gather flushes groups of 8 descriptors");

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// The gather is bound by memory latency per descriptor, so occupancy matters:
// it is held to 8 waves per SIMD (at most 64 VGPRs); at 76 VGPRs and 6 waves
// it ran 7 % slower than at 7.
template <bool kViews>
__global__ __launch_bounds__(kTaps) __attribute__((amdgpu_waves_per_eu(kViews ? 7 : 8))) void gather_kernel(GatherArgs a, const float4* __restrict__ pos_act,
                                                       const float4* __restrict__ fx, const float4* __restrict__ fy,
                                                       const float4* __restrict__ fz, const uint32_t* __restrict__ list,
                                                       void* __restrict__ feat_out, float* __restrict__ gamma,
                                                       float* __restrict__ pool) {
    // the descriptor row in LDS, stream k at k * kRowPitch: the pitch is 8
    // banks past a multiple of 32, so the three streams' segments start in
    // different banks for the pooling lanes
    constexpr int kRowPitch = kTaps + 8;
    __shared__ float sbuf[2][3 * kRowPitch];
    float* __restrict__ feat = (float*)feat_out;
    // a grid sized for n_items ends early.  Wave-uniform (readfirstlane), so the
    // loop does not wait at every item for this load, which on CDNA would also
    // wait for the previous item's stores (one counter for both).  Read first:
    // the grid covers the whole chunk, and most workgroups have no item, so they
    // leave before the per-tap set-up (uniform over the workgroup: no barrier
    // is split).  A capped grid striding over the groups instead measured
    // slower (0.598 against 0.580 ms at 1080p): the empty workgroups cost less
    // than the uneven split and the two registers the loop spills.
    const uint32_t n_items = __builtin_amdgcn_readfirstlane(a.n_dev ? *a.n_dev : a.n_items);
    if (blockIdx.x * kGatherItems >= n_items) return;
    const int i = threadIdx.x;
    const int mip = c_tap_mip[i];
    const float off = c_tap_offset[i];
    const float lut_v = c_tap_v[i];
    const int tm = mip - 1 > 0 ? mip - 1 : 0;
    const int tr = 128 >> tm;
    const vec3 st = i < kShellTaps ? v3(c_stencil[3 * i], c_stencil[3 * i + 1], c_stencil[3 * i + 2]) : v3(0, 0, 0);
    const float s_inv = i < kShellTaps ? c_tap_inv_len[i] : 0.0f;
    // The view's tables and light (with kViews, loaded when the view changes:
    // compacted items of one workgroup mostly share one), and the HG-table
    // factor that is the same for every descriptor of the view (TapTable): the
    // view-axis factor of a shell tap, the light-axis factor of a light tap.
    const float4* mip_cells = a.mips[mip];
    const float4* tr_cells = a.trs[tm];
    const float4* lut = a.lut;
    vec3 light = a.light, lx = a.lx;
    float alpha = a.alpha, g = a.g;
    const auto fixed_factor = [&]() {
        if (i > 0 && i < kShellTaps) return bilinear_cells(lut, kHgLutSize, kHgLutSize, c_tap_u_view[i], lut_v);
        if (i >= kShellTaps) return bilinear_cells(lut, kHgLutSize, kHgLutSize, dot(lx, light) * 0.5f + 0.5f, lut_v);
        return 0.0f;
    };
    float h_fixed = 0.0f;
    if constexpr (!kViews) h_fixed = fixed_factor();
    uint32_t cur_view = 0xffffffffu;
    // The vertex a descriptor belongs to (compacted list) and its record are
    // uniform over the block (scalar loads).  They are prefetched: the list
    // entry two items ahead, the record one item ahead, so an item's taps
    // start without waiting for either.
    const uint32_t item0 = blockIdx.x * kGatherItems;
    uint32_t src_next = 0u;
    float4 pa_n = make_float4(0.f, 0.f, 0.f, 0.f), x4_n = pa_n, y4_n = pa_n, z4_n = pa_n;
    if (item0 < n_items) {
        const uint32_t src0 = list ? list[item0] : item0;
        pa_n = pos_act[src0];
        x4_n = fx[src0];
        y4_n = fy[src0];
        z4_n = fz[src0];
        if (item0 + 1u < n_items) src_next = list ? list[item0 + 1u] : item0 + 1u;
    }
    for (int it = 0; it < kGatherItems; ++it) {
        const uint32_t item = item0 + (uint32_t)it;
        if (item >= n_items) break;  // uniform over the block
        const float4 pa = pa_n, x4 = x4_n, y4 = y4_n, z4 = z4_n;
        if (it + 1 < kGatherItems && item + 1u < n_items) {
            pa_n = pos_act[src_next];
            x4_n = fx[src_next];
            y4_n = fy[src_next];
            z4_n = fz[src_next];
            if (it + 2 < kGatherItems && item + 2u < n_items) src_next = list ? list[item + 2u] : item + 2u;
        }
        if constexpr (kViews) {
            const uint32_t view = __builtin_amdgcn_readfirstlane(__float_as_uint(y4.w) / a.view_items);
            if (view != cur_view) {
                const GatherArgs& va = a.views[view];
                mip_cells = va.mips[mip];
                tr_cells = va.trs[tm];
                lut = va.lut;
                light = va.light;
                lx = va.lx;
                alpha = va.alpha;
                g = va.g;
                h_fixed = fixed_factor();
                cur_view = view;
            }
        }
        const vec3 p = v3(pa.x, pa.y, pa.z);
        const vec3 X = v3(x4.x, x4.y, x4.z), Y = v3(y4.x, y4.y, y4.z), Z = v3(z4.x, z4.y, z4.z);
        float xd = 0.0f, xt = 0.0f, xh = 0.0f;
        if (pa.w != 0.0f) {
            vec3 pi;
            if (i < kShellTaps) {
                pi = p + (X * st.x + Y * st.y + Z * st.z) * off;
            } else {
                pi = p + light * off;
            }
            const float d = trilinear_cells(mip_cells, 256 >> mip, pi);
            xd = logf_(1.0f + alpha * d / 64.0f);
            float harg;
            if (i == 0) {
                harg = hg_phase(dot(X, light), g);
            } else {
                // the factor that varies: light axis for a shell tap, view axis for a light
                // tap (whose direction is the light's); the product commutes exactly
                const float uv = i < kShellTaps
                                     ? tap_u_light(dot(lx, X), dot(lx, Y), dot(lx, Z), st.x, st.y, st.z, s_inv)
                                     : dot(X, light) * 0.5f + 0.5f;
                harg = h_fixed * bilinear_cells(lut, kHgLutSize, kHgLutSize, uv, lut_v);
            }
            xh = logf_(harg + 1.0f);
            // a tap outside the box reads the TR mip where its light ray enters
            // the box (1 if it misses); branch-free: both cases read one cell
            const bool outside =
                pi.x < -0.5f || pi.y < -0.5f || pi.z < -0.5f || pi.x > 0.5f || pi.y > 0.5f || pi.z > 0.5f;
            const float o = ray_box_offset(pi, lx);
            const vec3 q = outside ? (pi + lx * o) + 0.5f : pi + 0.5f;
            const float tv = trilinear_clamp_cells(tr_cells, tr, q);
            xt = outside && o < 0.0f ? 1.0f : tv;
        }
        if (i == 0) gamma[item] = acosf_(dot(X, lx));  // after the lookups: stores delay later waits
        float* out = feat + (size_t)item * kFeatures;
        out[i] = xd;
        out[kTaps + i] = xt;
        out[2 * kTaps + i] = xh;
        if (pool) {
            float* sf = sbuf[it & 1];
            sf[i] = xd;
            sf[kRowPitch + i] = xt;
            sf[2 * kRowPitch + i] = xh;
            __syncthreads();
            if (i < 36) {
                // lane (block, stream) computes both statistics of its segment
                const int q = i, blk = q / 3, k = q % 3;
                const float* seg = sf + k * kRowPitch + c_seg_from[blk];
                const int n = c_seg_size[blk];  // 8, 16 or 32, per thread
                // 8 reads in flight per step (the loop bound differs across the
                // wave, so a plain loop would wait for each read); same order of sums
                float sum = 0.0f, mx = -10000.0f;
                for (int j0 = 0; j0 < n; j0 += 8) {
                    float v[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = seg[j0 + e];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        sum += v[e];
                        mx = fmax_(mx, v[e]);
                    }
                }
                pool[(size_t)item * 72 + i] = sum / (float)n;
                pool[(size_t)item * 72 + 36 + i] = mx;
            }
        }
    }
}

// ---------------------------------------------------------------- stage 2, tiled
// The gather of the MFMA renderers (wave-tiled BF16 descriptors and pooled
// statistics), with the same arithmetic per feature as gather_kernel, so every
// feature is the BF16 rounding of the FP32 gather's.  What bounds it is the
// caches' misses: with every lookup made to hit (a timing copy that reads
// cell 0) it took 0.33 ms against 0.55 ms at 1080p; the density lookups'
// misses cost the most (0.14 ms), the transmittance and HG-table lookups' 0.08
// ms each.  Against gather_kernel's tiled path (0.58 ms) it issues a quarter
// fewer vector instructions:
//   - values shared by every tap of a descriptor (tap 0's phase, the light
//     taps' table coordinate, the light axis in the descriptor's frame,
//     gamma) are computed once per workgroup, one lane per descriptor, and
//     read from LDS with the descriptor's record;
//   - shell and light taps, and tap 0, take the same branch-free path (the
//     old one ran both sides of three branches in some waves);
//   - each tap's HG-table row (its v is fixed) is placed once, not per lookup;
//   - clamps, box tests and the ray-box interval use min3 / max3 / med3,
//     which equal the compare-and-select forms on these inputs (no NaN; an
//     exact tie gives the same value either way, and a zero from exact
//     cancellation is +0, so no +0 / -0 tie occurs);
//   - cell indices use 24-bit multiply-adds (indices stay below 2^24 before
//     the last multiply, products below 2^32).
struct GatherRec {
    float4 p;  // position, active flag
    float4 x;  // frame X, tap 0's HG phase
    float4 y;  // frame Y, the light taps' HG-table u
    float4 z;  // frame Z, view index (bits)
    float4 l;  // light axis . (X, Y, Z) for the shell taps' HG-table u
};

// Per-thread constants of a view: its tables and the values every descriptor shares.
struct GatherView {
    const float4* mip_cells;
    const float4* tr_cells;
    const float4* lut;
    vec3 light, lx, id, tf;  // id = inv3(lx), tf = 0.4999 * id (RayBoxOffset's per-direction terms)
    float alpha, g, h_fixed;
};

__device__ __forceinline__ uint32_t cell_index24(uint32_t c, uint32_t ux, uint32_t uy, uint32_t uz) {
    return mad_u24(mad_u24(ux, c, uy), c, uz);
}

// 7 waves per SIMD (73 VGPRs): at 8 (64 VGPRs) about ten values spill to
// scratch.  The multi-view form keeps the view's constants for a change of
// view as well and spills 52 bytes per lane at 7, so it runs at 6.
template <bool kCells16, bool kViews>
__global__ __launch_bounds__(kTaps) __attribute__((amdgpu_waves_per_eu(kViews ? 5 : 7))) void gather_tiled_kernel(
    GatherArgs a, const float4* __restrict__ pos_act, const float4* __restrict__ fx, const float4* __restrict__ fy,
    const float4* __restrict__ fz, const uint32_t* __restrict__ list, void* __restrict__ feat_out,
    float* __restrict__ gamma, float* __restrict__ pool) {
    // the descriptor row in LDS, stream k at k * kRowPitch: the pitch is 8
    // banks past a multiple of 32, so the three streams' segments start in
    // different banks for the pooling lanes
    constexpr int kRowPitch = kTaps + 8;
    __shared__ float sbuf[2][3 * kRowPitch];
    // BF16 chunks of kStageItems consecutive descriptors, flushed as whole
    // 128-byte lines.  Staged chunk q of item j sits at stage[j * kStagePitch + q]:
    // the chunk writes of one item are consecutive 16-byte words, and the
    // flush's 8 lanes of one chunk read 8 rows whose pitch of 73 words spreads
    // them over all banks
    constexpr int kStageItems = 8, kChunks = 3 * 2 * kStreamChunks, kStagePitch = kChunks + 1;
    __shared__ bf16x8_t stage[kStagePitch * kStageItems];
    __shared__ GatherRec srec[kGatherItems];
    static_assert(!(kViews && kCells16), "the multi-view gather reads the FP32 cells");
    const uint32_t n_items = __builtin_amdgcn_readfirstlane(a.n_dev ? *a.n_dev : a.n_items);
    if (blockIdx.x * kGatherItems >= n_items) return;  // uniform: before any barrier
    const int i = threadIdx.x;
    const uint32_t item0 = blockIdx.x * kGatherItems;

    // records of this workgroup's descriptors, one lane each
    if (i < kGatherItems && item0 + (uint32_t)i < n_items) {
        const uint32_t item = item0 + (uint32_t)i;
        const uint32_t src = list ? list[item] : item;
        const float4 pa = pos_act[src], x4 = fx[src], y4 = fy[src], z4 = fz[src];
        const GatherArgs* va = &a;
        uint32_t view = 0u;
        if constexpr (kViews) {
            view = __float_as_uint(y4.w) / a.view_items;
            va = &a.views[view];
        }
        const vec3 X = v3(x4.x, x4.y, x4.z), Y = v3(y4.x, y4.y, y4.z), Z = v3(z4.x, z4.y, z4.z);
        const vec3 light = va->light, lx = va->lx;
        const float xl = dot(X, light);
        GatherRec r;
        r.p = pa;
        r.x = make_float4(X.x, X.y, X.z, hg_phase(xl, va->g));
        r.y = make_float4(Y.x, Y.y, Y.z, xl * 0.5f + 0.5f);
        r.z = make_float4(Z.x, Z.y, Z.z, __uint_as_float(view));
        r.l = make_float4(dot(lx, X), dot(lx, Y), dot(lx, Z), 0.0f);
        srec[i] = r;
        gamma[item] = acosf_(dot(X, lx));
    }

    const int mip = c_tap_mip[i];
    const float off = c_tap_offset[i];
    const float lut_v = c_tap_v[i];
    const int tm = mip - 1 > 0 ? mip - 1 : 0;
    const int tr = 128 >> tm;
    const int dres = 256 >> mip;
    const bool shell = i < kShellTaps;
    const vec3 st = shell ? v3(c_stencil[3 * i], c_stencil[3 * i + 1], c_stencil[3 * i + 2]) : v3(0, 0, 0);
    const float s_inv = shell ? c_tap_inv_len[i] : 0.0f;
    const float u_view = shell ? c_tap_u_view[i] : 0.0f;
    const float fr_d = (float)dres, fr_t = (float)tr;
    const uint32_t c_d = (uint32_t)dres + 1u, c_t = (uint32_t)tr + 2u;
    // the HG-table row of the tap (v is fixed per tap), as bilinear_cells places it
    const float lty = __builtin_amdgcn_fmed3f(lut_v * (float)kHgLutSize - 0.5f, -1.0f, (float)kHgLutSize);
    const float lfy = floorf(lty), l_wy = lty - lfy;
    const uint32_t l_row = __umul24((uint32_t)(lfy + 1.0f), (uint32_t)kHgLutSize + 2u);
    const uint4* mips16 = kCells16 ? a.mips16 + mip16_offset(mip) : nullptr;
    const uint4* trs16 = kCells16 ? a.trs16 + tr16_offset(tm) : nullptr;

    GatherView gv;
    const auto set_view = [&](const GatherArgs& va) {
        gv.mip_cells = va.mips[mip];
        gv.tr_cells = va.trs[tm];
        gv.lut = va.lut;
        gv.light = va.light;
        gv.lx = va.lx;
        gv.id = va.id;
        gv.tf = va.tf;
        gv.alpha = va.alpha;
        gv.g = va.g;
        // the HG-table factor that is the same for every descriptor of the view:
        // the view-axis factor of a shell tap, the light-axis factor of a light tap
        gv.h_fixed = 0.0f;
        if (i > 0 && shell) gv.h_fixed = bilinear_cells(gv.lut, kHgLutSize, kHgLutSize, u_view, lut_v);
        if (!shell) gv.h_fixed = bilinear_cells(gv.lut, kHgLutSize, kHgLutSize, dot(va.lx, va.light) * 0.5f + 0.5f, lut_v);
    };
    uint32_t cur_view = 0xffffffffu;
    if constexpr (!kViews) set_view(a);
    __syncthreads();

    for (int it = 0; it < kGatherItems; ++it) {
        const uint32_t item = item0 + (uint32_t)it;
        if (item >= n_items) break;  // uniform over the block
        const GatherRec r = srec[it];
        if constexpr (kViews) {
            const uint32_t view = __builtin_amdgcn_readfirstlane(__float_as_uint(r.z.w));
            if (view != cur_view) {
                set_view(a.views[view]);
                cur_view = view;
            }
        }
        float xd = 0.0f, xt = 0.0f, xh = 0.0f;
        if (r.p.w != 0.0f) {  // uniform over the block
            const vec3 p = v3(r.p.x, r.p.y, r.p.z);
            const vec3 X = v3(r.x.x, r.x.y, r.x.z), Y = v3(r.y.x, r.y.y, r.y.z), Z = v3(r.z.x, r.z.y, r.z.z);
            const vec3 sd = X * st.x + Y * st.y + Z * st.z;
            const vec3 dir = shell ? sd : gv.light;
            const vec3 pi = p + dir * off;

            // density: border cells of mip `mip`
            const float dtx = (pi.x + 0.5f) * fr_d - 0.5f;
            const float dty = (pi.y + 0.5f) * fr_d - 0.5f;
            const float dtz = (pi.z + 0.5f) * fr_d - 0.5f;
            const bool din = __builtin_fminf(__builtin_fminf(dtx, dty), dtz) >= -1.0f &&
                             __builtin_fmaxf(__builtin_fmaxf(dtx, dty), dtz) < fr_d;
            const float dfx = din ? floorf(dtx) : -1.0f, dfy = din ? floorf(dty) : -1.0f,
                        dfz = din ? floorf(dtz) : -1.0f;
            // cell coordinates + 1 in [0, res]: the +1 is exact in FP32
            const uint32_t dci = cell_index24(c_d, (uint32_t)(dfx + 1.0f), (uint32_t)(dfy + 1.0f),
                                              (uint32_t)(dfz + 1.0f));

            // HG table: u of the tap's direction against the light axis (shell) or
            // of the light against the view axis (light taps); tap 0 uses its phase
            const float u = shell ? tap_u_light(r.l.x, r.l.y, r.l.z, st.x, st.y, st.z, s_inv) : r.y.w;
            const float ltx = __builtin_amdgcn_fmed3f(u * (float)kHgLutSize - 0.5f, -1.0f, (float)kHgLutSize);
            const float lfx = floorf(ltx);
            const uint32_t lci = l_row + (uint32_t)(lfx + 1.0f);

            // transmittance: a tap outside the box reads the TR mip where its light
            // ray enters the box (1 if it misses)
            const bool outside = __builtin_fminf(__builtin_fminf(pi.x, pi.y), pi.z) < -0.5f ||
                                 __builtin_fmaxf(__builtin_fmaxf(pi.x, pi.y), pi.z) > 0.5f;
            const vec3 to_axis = (-pi) * gv.id;
            const vec3 d0 = to_axis + gv.tf, d1 = to_axis - gv.tf;
            const float tmi = __builtin_fmaxf(
                __builtin_fmaxf(__builtin_fminf(d0.x, d1.x), __builtin_fminf(d0.y, d1.y)), __builtin_fminf(d0.z, d1.z));
            const float tma = __builtin_fminf(
                __builtin_fminf(__builtin_fmaxf(d0.x, d1.x), __builtin_fmaxf(d0.y, d1.y)), __builtin_fmaxf(d0.z, d1.z));
            const float o = tma >= tmi ? __builtin_fmaxf(tmi, 0.0f) : -1.0f;
            const vec3 q = outside ? (pi + gv.lx * o) + 0.5f : pi + 0.5f;
            const float ttx = __builtin_amdgcn_fmed3f(q.x * fr_t - 0.5f, -1.0f, fr_t);
            const float tty = __builtin_amdgcn_fmed3f(q.y * fr_t - 0.5f, -1.0f, fr_t);
            const float ttz = __builtin_amdgcn_fmed3f(q.z * fr_t - 0.5f, -1.0f, fr_t);
            const float tfx = floorf(ttx), tfy = floorf(tty), tfz = floorf(ttz);
            const uint32_t tci = cell_index24(c_t, (uint32_t)(tfx + 1.0f), (uint32_t)(tfy + 1.0f),
                                              (uint32_t)(tfz + 1.0f));

            // The three lookups one after another: density, HG table, transmittance.
            // With all five loads of a tap issued at once (more lines in flight per
            // CU) the kernel ran 0.72 ms against 0.56 ms at 1080p: it is bound by
            // the caches' misses, not by instruction issue.  The scheduling
            // barriers keep the compiler from hoisting the next lookup's loads.
            float d, tv;
            if constexpr (kCells16) {
                d = blend_cell16(mips16[dci], dtx - dfx, dty - dfy, dtz - dfz);
            } else {
                const float4* dc = gv.mip_cells + 2u * dci;
                const float4 dlo = dc[0], dhi = dc[1];
                const float dwx = dtx - dfx, dwy = dty - dfy, dwz = dtz - dfz;
                d = lerp(lerp(lerp(dlo.x, dlo.y, dwz), lerp(dlo.z, dlo.w, dwz), dwy),
                         lerp(lerp(dhi.x, dhi.y, dwz), lerp(dhi.z, dhi.w, dwz), dwy), dwx);
            }
            d = din ? d : 0.0f;
            xd = logf_(1.0f + gv.alpha * d / 64.0f);
            __builtin_amdgcn_sched_barrier(0);
            const float4 lq = gv.lut[lci];
            const float lwx = ltx - lfx;
            const float hl = lerp(lerp(lq.x, lq.y, lwx), lerp(lq.z, lq.w, lwx), l_wy);
            const float harg = i == 0 ? r.x.w : gv.h_fixed * hl;
            xh = logf_(harg + 1.0f);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (kCells16) {
                tv = blend_cell16(trs16[tci], ttx - tfx, tty - tfy, ttz - tfz);
            } else {
                const float4* tc = gv.tr_cells + 2u * tci;
                const float4 tlo = tc[0], thi = tc[1];
                const float twx = ttx - tfx, twy = tty - tfy, twz = ttz - tfz;
                tv = lerp(lerp(lerp(tlo.x, tlo.y, twz), lerp(tlo.z, tlo.w, twz), twy),
                          lerp(lerp(thi.x, thi.y, twz), lerp(thi.z, thi.w, twz), twy), twx);
            }
            xt = outside && o < 0.0f ? 1.0f : tv;
        }
        float* sf = sbuf[it & 1];
        sf[i] = xd;
        sf[kRowPitch + i] = xt;
        sf[2 * kRowPitch + i] = xh;
        __syncthreads();
        if (i >= 96 && i < 96 + kChunks) {
            // chunk q = (stream k, 16-feature chunk c, lane group kg)
            const int q = i - 96, k = q / (2 * kStreamChunks), c = (q >> 1) % kStreamChunks, kg = q & 1;
            const float4* src = (const float4*)(sf + k * kRowPitch + 16 * c + 4 * kg);
            const float4 lo = src[0], hi = src[2];
            bf16x8_t v;
            v[0] = (__bf16)lo.x;
            v[1] = (__bf16)lo.y;
            v[2] = (__bf16)lo.z;
            v[3] = (__bf16)lo.w;
            v[4] = (__bf16)hi.x;
            v[5] = (__bf16)hi.y;
            v[6] = (__bf16)hi.z;
            v[7] = (__bf16)hi.w;
            stage[(it & (kStageItems - 1)) * kStagePitch + q] = v;
        }
        if (i < 36) {
            // lane (block, stream) computes both statistics of its segment, in order
            const int blk = i / 3, k = i % 3;
            const float4* seg = (const float4*)(sf + k * kRowPitch + c_seg_from[blk]);
            const int n = c_seg_size[blk];  // 8, 16 or 32, per thread
            float sum = 0.0f, mx = -10000.0f;
            for (int j0 = 0; j0 < n; j0 += 8) {
                const float4 v0 = seg[j0 / 4], v1 = seg[j0 / 4 + 1];
                const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    sum += v[e];
                    mx = fmax_(mx, v[e]);
                }
            }
            pool[(size_t)item * 72 + i] = sum / (float)n;
            pool[(size_t)item * 72 + 36 + i] = mx;
        }
        if ((it & (kStageItems - 1)) == kStageItems - 1 || item + 1 >= n_items) {
            // uniform over the block; items start at multiples of kStageItems
            __syncthreads();
            const int cnt = (it & (kStageItems - 1)) + 1;
            const uint32_t base = item + 1u - (uint32_t)cnt;
            for (int idx = i; idx < kChunks * kStageItems; idx += kTaps) {
                const int q = idx / kStageItems, j = idx % kStageItems;
                if (j < cnt) {
                    const int k = q / (2 * kStreamChunks), c = (q >> 1) % kStreamChunks, kg = q & 1;
                    ((bf16x8_t*)feat_out)[tile_chunk(base + (uint32_t)j, k, c, kg)] = stage[j * kStagePitch + q];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- stage 3 (FP32)
__global__ __launch_bounds__(64) void mlp_fp32_kernel(const float* __restrict__ w, const float* __restrict__ feat,
                                                      const float* __restrict__ gamma, float g, vec3 srp,
                                                      uint32_t n_items, const uint32_t* __restrict__ n_dev,
                                                      float* __restrict__ pred) {
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (n_dev ? *n_dev : n_items)) return;
    float o[3];
    mlp_forward_fp32(w, feat + (size_t)item * kFeatures, gamma[item], g, srp, o);
    pred[3 * item] = o[0];
    pred[3 * item + 1] = o[1];
    pred[3 * item + 2] = o[2];
}

// ---------------------------------------------------------------- stage 4
// kViews: the light colour and albedo are those of the slot's view.
template <bool kViews>
__global__ __launch_bounds__(256) void combine_kernel(const float4* __restrict__ pos_act,
                                                      const float* __restrict__ single_in,
                                                      const float4* __restrict__ fx,
                                                      const float4* __restrict__ fy, const float4* __restrict__ fz,
                                                      const uint32_t* __restrict__ slot_of,
                                                      const float* __restrict__ pred, vec3 light_color,
                                                      vec3 scatter_rate, const NNArgs* __restrict__ views,
                                                      uint32_t view_items, uint32_t n_items,
                                                      float4* __restrict__ out, SampleFold fold) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const uint32_t slot = __float_as_uint(fy[i].w);
    if (slot == 0xffffffffu) return;
    float4 o;
    if (pos_act[i].w != 0.0f) {
        if constexpr (kViews) {
            const NNArgs& va = views[slot / view_items];
            light_color = va.light_color;
            scatter_rate = va.scatter_rate;
        }
        const float single = single_in[i];
        const uint32_t k = slot_of[i];  // the vertex's compacted descriptor
        const vec3 p = v3(pred[3 * k] + single, pred[3 * k + 1] + single, pred[3 * k + 2] + single);
        const vec3 r = p * light_color * scatter_rate;
        o = make_float4(r.x, r.y, r.z, fz[i].w);
    } else {
        const float4 sky = fx[i];  // the environment radiance stored by the first vertex
        o = make_float4(sky.x, sky.y, sky.z, -1.0f);
    }
    if (!kViews && fold.target) {
        // one sample per pixel: accumulate_kernel's fold of this pixel's one sample
        const vec3 c = vmax(v3(0, 0, 0), v3(o.x, o.y, o.z));
        float* t = fold.target + 3 * (size_t)slot;
        const vec3 tg = lerp(v3(t[0], t[1], t[2]), c, fold.rate);
        t[0] = tg.x;
        t[1] = tg.y;
        t[2] = tg.z;
        if (fold.sum64) {
            double* s = fold.sum64 + 3 * (size_t)slot;
            s[0] += (double)c.x;
            s[1] += (double)c.y;
            s[2] += (double)c.z;
        }
    } else {
        out[slot] = o;
    }
}

// Active first vertices -> a dense list (wave ballots, one atomic per workgroup).
// The list order depends on scheduling; every descriptor's result does not.
constexpr int kCompactBlock = 1024;
template <bool kViews>
__global__ __launch_bounds__(kCompactBlock) void compact_kernel(NNArgs a0, const float4* __restrict__ pos_act,
                                                                const float4* __restrict__ fy, uint32_t n,
                                                                uint32_t* __restrict__ list,
                                                                uint32_t* __restrict__ slot_of,
                                                                uint32_t* __restrict__ count,
                                                                uint32_t* __restrict__ buckets) {
    // one atomic per workgroup (a wave ballot per wave, a scan of the wave
    // counts in LDS): the counter's atomics serialise.  The listed vertices
    // are also counted into the shadow stage's length buckets.
    __shared__ uint32_t wave_off[kCompactBlock / 64];
    __shared__ uint32_t block_base;
    __shared__ uint32_t hist[kShadowBuckets];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < kShadowBuckets) hist[threadIdx.x] = 0u;
    const bool act = i < n && pos_act[i].w != 0.0f;
    const unsigned long long m = __ballot(act);
    const uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
    if (lane == 0u) wave_off[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0u;
        for (int q = 0; q < kCompactBlock / 64; ++q) {
            const uint32_t c = wave_off[q];
            wave_off[q] = tot;
            tot += c;
        }
        block_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = block_base + wave_off[wv];
    if (i < n) {
        if (act) {
            const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            list[k] = i;
            slot_of[i] = k;
            atomicAdd(&hist[shadow_bucket<kViews>(a0, pos_act, fy, i)], 1u);
        } else {
            slot_of[i] = 0xffffffffu;
        }
    }
    __syncthreads();
    if (threadIdx.x < kShadowBuckets && hist[threadIdx.x]) atomicAdd(&buckets[threadIdx.x], hist[threadIdx.x]);
}


__global__ void clear_counters_kernel(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

static vec3 V(const mrp_vec3& a) { return v3(a.x, a.y, a.z); }



static bool g_tables_ready[64] = {};
static std::mutex g_tables_mutex;

static int upload_tap_tables(int device) {
    std::lock_guard<std::mutex> lock(g_tables_mutex);  // volumes may render from several threads
    if (device >= 0 && device < 64 && g_tables_ready[device]) return MRP_OK;
    TapTable t;
    descriptor_schedule(&t);
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tap_offset), t.offset, sizeof(t.offset)));
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tap_mip), t.mip, sizeof(t.mip)));
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tap_v), t.lut_v, sizeof(t.lut_v)));
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_stencil), &t.stencil[0][0], sizeof(t.stencil)));
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tap_inv_len), t.inv_len, sizeof(t.inv_len)));
    MRP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tap_u_view), t.u_view, sizeof(t.u_view)));
    if (device >= 0 && device < 64) g_tables_ready[device] = true;
    return MRP_OK;
}

// Per-item scratch of the MRPNN stages (bytes): 4 float4 first-vertex
// records, 576 FP32 features, gamma, 3 predictions.
constexpr size_t kNNItemBytes = 4 * 16 + kFeatures * 4 + 72 * 4 + 4 + 12 + 8 + 4 + 4;  // + list, slot, stream position, single

// The MRPNN stages' scratch lives in the volume (nn_scratch), so volumes on
// different streams can render concurrently.
static int ensure_nn_scratch(Volume* v, size_t bytes, void** out) {
    if (v->nn_scratch_bytes < bytes) {
        if (v->nn_scratch) {
            MRP_HIP(hipStreamSynchronize(v->stream));
            MRP_HIP(hipFree(v->nn_scratch));
        }
        v->nn_scratch = nullptr;
        v->nn_scratch_bytes = 0;
        if (hipMalloc(&v->nn_scratch, bytes) != hipSuccess) return fail(MRP_E_NOMEM, "MRPNN scratch allocation failed");
        v->nn_scratch_bytes = bytes;
    }
    *out = v->nn_scratch;
    return MRP_OK;
}

static vec3 srp_of(vec3 scatter_rate) {
    const vec3 sr = v3(scatter_rate.x / 1.001f, scatter_rate.y / 1.001f, scatter_rate.z / 1.001f);
    return v3((float)std::pow((double)sr.x, 4.0), (float)std::pow((double)sr.y, 4.0), (float)std::pow((double)sr.z, 4.0));
}

// The tiled (BF16/FP8 network) gather on the volume's cell format.
static int launch_gather_tiled(Volume* v, const GatherArgs& ga, unsigned blocks, const float4* pos_act,
                               const float4* fx, const float4* fy, const float4* fz, const uint32_t* list,
                               void* tiles, float* gamma, float* pool) {
    if (v->gather_cells16)
        gather_tiled_kernel<true, false><<<blocks, kTaps, 0, v->stream>>>(ga, pos_act, fx, fy, fz, list, tiles, gamma, pool);
    else
        gather_tiled_kernel<false, false><<<blocks, kTaps, 0, v->stream>>>(ga, pos_act, fx, fy, fz, list, tiles, gamma, pool);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

// Brings the volume's per-light and per-g tables (TR mips, HG table) up to
// date on its stream.
static int prepare_tables(Volume* v, const mrp_render_params* p) {
    int rc = upload_tap_tables(v->device);
    if (rc) return rc;
    rc = ensure_tr(v, p->light_dir, p->alpha);
    if (rc) return rc;
    return ensure_hg_lut(v, p->g, p->seed);
}

// Checks a render's network settings, then prepare_tables.
static int prepare_volume(Volume* v, const mrp_render_params* p, const void** packed) {
    if (!v->nn_weights) return fail(MRP_E_STATE, "no network set (mrp_volume_set_network)");
    if (p->mlp_precision < 0 || p->mlp_precision > 2)
        return fail(MRP_E_INVALID, "mlp_precision must be 0 (FP32), 1 (BF16) or 2 (FP8)");
    *packed = p->mlp_precision == 2 ? v->nn_packed_fp8 : v->nn_packed;
    if (p->mlp_precision != 0 && !*packed) return fail(MRP_E_STATE, "network has no MFMA packing");
    return prepare_tables(v, p);
}

// The stage records of one image: samples from s_begin of the pixels `map`
// selects, on volume v.
static void image_args(const Volume* v, const mrp_render_params* p, PixelMap map, int s_begin, NNArgs* a,
                       GatherArgs* ga) {
    const int W = p->width, H = p->height;
    *a = NNArgs{};
    a->cells = v->cells;
    a->quads = v->quads;
    a->qg = quad_geom(v->res);
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
    a->light = normalize(V(p->light_dir));
    a->lx = normalize(a->light);  // first axis of the light frame (GetMatrixFromNormal normalises again)
    a->light_color = V(p->light_color);
    a->env = env_map(v, V(p->sky));
    a->scatter_rate = V(p->scatter_rate);
    a->width = W;
    a->height = H;
    a->checker = map.checker;
    a->rows_up = map.rows_up;
    a->ray_ori = map.ray_ori;
    a->ray_dir = map.ray_dir;
    a->wl = map.checker < 0 ? W : (W + 1) / 2;
    a->seed_lo = (uint32_t)p->seed;
    a->seed_hi = (uint32_t)(p->seed >> 32);
    a->spp_begin = s_begin;
    a->counters = v->counters;
    a->views = nullptr;
    a->view_items = 0u;
    a->occ_words = 0u;

    *ga = GatherArgs{};
    for (int m = 0; m < kNumMips; ++m) ga->mips[m] = v->mip_cells[m];
    for (int m = 0; m < kNumTrMips; ++m) ga->trs[m] = v->tr_cells[m];
    ga->lut = v->lut_cells;
    ga->mips16 = v->mip_cells16;
    ga->trs16 = v->tr_cells16;
    ga->light = a->light;
    ga->lx = a->lx;
    ga->id = inv3(ga->lx);
    ga->tf = v3(0.4999f, 0.4999f, 0.4999f) * ga->id;
    ga->alpha = p->alpha;
    ga->g = p->g;
    ga->views = nullptr;
    ga->view_items = 0u;
}

// First vertex -> compaction -> shadow rays -> gather -> network -> combine
// over `total` items in chunks of at most 4M, on v's stream and scratch; the
// results go to out[slot].  With a.views set the items run over several
// images (the kernels' kViews forms); the network settings (precision,
// weights, g, srp) are shared by all of them.
static int run_stages(Volume* v, NNArgs a, GatherArgs ga, int precision, const void* packed, const float* weights,
                      float g, vec3 srp, size_t total, float4* out, NNTimes* times, uint64_t* n_active,
                      bool defer_last, SampleFold fold = SampleFold{}) {
    const bool views = a.views != nullptr;
    const size_t chunk_items = std::max<size_t>(256, std::min<size_t>((size_t)1 << 22, total));
    static_assert(2 * kTileBytes == (size_t)32 * kFeatures * sizeof(float), "tiled BF16 descriptors fit the FP32 rows");
    void* scratch = nullptr;
    int rc = ensure_nn_scratch(v, chunk_items * kNNItemBytes + 4096, &scratch);
    if (rc) return rc;
    char* sp = (char*)scratch;
    float4* pos_act = (float4*)sp;
    float4* fx = pos_act + chunk_items;
    float4* fy = fx + chunk_items;
    float4* fz = fy + chunk_items;
    float* feat = (float*)(fz + chunk_items);
    float* pool = feat + chunk_items * kFeatures;
    float* gam = pool + chunk_items * 72;
    float* pred = gam + chunk_items;
    uint32_t* list = (uint32_t*)(pred + 3 * chunk_items);
    uint32_t* slot_of = list + chunk_items;
    uint32_t* rng_pos = slot_of + chunk_items;
    float* single = (float*)(rng_pos + chunk_items);
    uint32_t* n_active_dev = (uint32_t*)v->queues;  // the tracer's queue counters are free during MRPNN frames

    // Stage boundaries: first vertex | compaction | shadow rays | gather | network | combine.
    // the volume's events, created once (creating them per frame costs host time between frames)
    for (hipEvent_t& e : v->nn_events)
        if (!e) MRP_HIP(hipEventCreate(&e));
    const hipEvent_t t0 = v->nn_events[0], tc = v->nn_events[1], ts0 = v->nn_events[2], ts1 = v->nn_events[3],
                     t2 = v->nn_events[4], t3 = v->nn_events[5], tj = v->nn_events[6];
    for (size_t i0 = 0; i0 < total && rc == MRP_OK; i0 += chunk_items) {
        const uint32_t cnt = (uint32_t)std::min(chunk_items, total - i0);
        a.item0 = (uint32_t)i0;
        a.n_items = cnt;
        const unsigned blocks = (cnt + 255) / 256;
        // the shadow stage's bucket counts and offsets (2 * kShadowBuckets words after the items)
        uint32_t* bucket_counts = (uint32_t*)(sp + chunk_items * kNNItemBytes);
        MRP_HIP(hipEventRecord(t0, v->stream));
        if (views)
            first_vertex_kernel<true><<<blocks, 256, a.occ_words * sizeof(uint32_t), v->stream>>>(
                a, pos_act, fx, fy, fz, rng_pos, n_active_dev, bucket_counts);
        else
            first_vertex_kernel<false><<<blocks, 256, occupancy_words(a.res) * sizeof(uint32_t), v->stream>>>(
                a, pos_act, fx, fy, fz, rng_pos, n_active_dev, bucket_counts);
        MRP_HIP(hipGetLastError());
        MRP_HIP(hipEventRecord(tc, v->stream));
        if (views)
            compact_kernel<true><<<(cnt + kCompactBlock - 1) / kCompactBlock, kCompactBlock, 0, v->stream>>>(
                a, pos_act, fy, cnt, list, slot_of, n_active_dev, bucket_counts);
        else
            compact_kernel<false><<<(cnt + kCompactBlock - 1) / kCompactBlock, kCompactBlock, 0, v->stream>>>(
                a, pos_act, fy, cnt, list, slot_of, n_active_dev, bucket_counts);
        MRP_HIP(hipGetLastError());
        // No read-back here: the shadow rays, the gather and the network read the
        // count on the device, with grids sized for the whole chunk.
        MRP_HIP(hipEventRecord(ts0, v->stream));
        // shadow rays in buckets of their length toward the light (the gather
        // keeps the compaction's pixel-coherent order); the order lives in the
        // descriptor region, which the gather fills only afterwards
        uint32_t* sorder = (uint32_t*)feat;
        if (views) {
            shadow_scatter_kernel<true><<<blocks, 256, 0, v->stream>>>(a, list, n_active_dev, pos_act, fy,
                                                                       bucket_counts, bucket_counts + kShadowBuckets,
                                                                       sorder);
            shadow_kernel<true><<<blocks, 256, 0, v->stream>>>(a, sorder, n_active_dev, pos_act, fy, rng_pos, fx,
                                                               single);
        } else {
            shadow_scatter_kernel<false><<<blocks, 256, 0, v->stream>>>(a, list, n_active_dev, pos_act, fy,
                                                                        bucket_counts, bucket_counts + kShadowBuckets,
                                                                        sorder);
            shadow_kernel<false><<<blocks, 256, occupancy_words(a.res) * sizeof(uint32_t), v->stream>>>(
                a, sorder, n_active_dev, pos_act, fy, rng_pos, fx,
                                                                single);
        }
        MRP_HIP(hipGetLastError());
        MRP_HIP(hipEventRecord(ts1, v->stream));
        ga.n_items = cnt;
        ga.n_dev = n_active_dev;
        {
            // the MFMA networks read wave-tiled BF16 descriptors (same region: ceil(n/32) tiles
            // of kTileBytes fit in n FP32 rows for n >= 32); the FP32 network reads FP32 rows
            const unsigned gblocks = (cnt + kGatherItems - 1) / kGatherItems;  // groups past the count exit at once
            if (views && precision == 0)
                gather_kernel<true><<<gblocks, kTaps, 0, v->stream>>>(ga, pos_act, fx, fy, fz, list, feat,
                                                                                   gam, pool);
            else if (views)
                gather_tiled_kernel<false, true><<<gblocks, kTaps, 0, v->stream>>>(ga, pos_act, fx, fy, fz, list, feat,
                                                                                  gam, pool);
            else if (precision == 0)
                gather_kernel<false><<<gblocks, kTaps, 0, v->stream>>>(ga, pos_act, fx, fy, fz, list,
                                                                                    feat, gam, pool);
            else
                rc = launch_gather_tiled(v, ga, gblocks, pos_act, fx, fy, fz, list, feat, gam, pool);
            MRP_HIP(hipGetLastError());
            if (rc) return rc;
        }
        MRP_HIP(hipEventRecord(t2, v->stream));
        if (precision == 0) {
            mlp_fp32_kernel<<<(cnt + 63) / 64, 64, 0, v->stream>>>(weights, feat, gam, g, srp, cnt, n_active_dev, pred);
            MRP_HIP(hipGetLastError());
        } else {
            rc = launch_mlp(packed, precision - 1, nullptr, feat, pool, gam, g, srp, cnt, n_active_dev, pred,
                            v->stream);
        }
        MRP_HIP(hipEventRecord(t3, v->stream));
        if (rc == MRP_OK) {
            if (views)
                combine_kernel<true><<<blocks, 256, 0, v->stream>>>(pos_act, single, fx, fy, fz, slot_of, pred,
                                                                    a.light_color, a.scatter_rate, a.views,
                                                                    a.view_items, cnt, out, SampleFold{});
            else
                combine_kernel<false><<<blocks, 256, 0, v->stream>>>(pos_act, single, fx, fy, fz, slot_of, pred,
                                                                     a.light_color, a.scatter_rate, nullptr, 0u, cnt,
                                                                     out, fold);
            MRP_HIP(hipGetLastError());
        }
        // the count into pinned host memory (a copy into pageable memory would block)
        MRP_HIP(hipMemcpyAsync(v->host_words, n_active_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, v->stream));
        if (rc != MRP_OK) break;
        v->nn_pending = true;
        if (defer_last && i0 + cnt >= total) break;  // the caller synchronises, then mrpnn_finish
        MRP_HIP(hipEventRecord(tj, v->stream));
        MRP_HIP(hipEventSynchronize(tj));
        rc = mrpnn_finish(v, times, n_active);
    }
    return rc;
}

// First vertex -> gather -> network -> combine for samples [s_begin, s_begin+n)
// of the pixels `map` selects, in chunks of at most 4M vertices.
int mrpnn_samples(Volume* v, const mrp_render_params* p, PixelMap map, int s_begin, int n, NNTimes* times,
                  uint64_t* n_active, bool defer_last, SampleFold fold) {
    const void* packed = nullptr;
    int rc = prepare_volume(v, p, &packed);
    if (rc) return rc;
    NNArgs a;
    GatherArgs ga;
    image_args(v, p, map, s_begin, &a, &ga);
    const size_t total = (size_t)n * (size_t)a.wl * (size_t)p->height;
    if (fold.target && (n != 1 || map.checker >= 0 || a.wl != p->width))
        return fail(MRP_E_INVALID, "a fold into the running mean needs one sample of every pixel");
    return run_stages(v, a, ga, p->mlp_precision, packed, v->nn_weights, p->g, srp_of(a.scatter_rate), total,
                      (float4*)v->samples, times, n_active, defer_last, fold);
}

size_t mrpnn_view_table_bytes(int n_views) { return (size_t)n_views * (sizeof(NNArgs) + sizeof(GatherArgs)); }

// The MRPNN samples [0, spp) of n_views images in one pass of the stages
// (view k: volume vols[k], parameters ps[k]): out[(k * spp + s) * P + pixel].
// The network of vols[0] and its settings (precision, g, albedo) serve every
// view; views that share a volume share its light and alpha (its TR mips).
// tables_dev holds n_views NNArgs then n_views GatherArgs.
int mrpnn_views(Volume* const* vols, const mrp_render_params* ps, int n_views, int spp, void* tables_dev,
                float4* out, NNTimes* times, uint64_t* n_active) {
    Volume* host = vols[0];
    const mrp_render_params& p0 = ps[0];
    const size_t P = (size_t)p0.width * p0.height;
    if (P % 64 != 0) return fail(MRP_E_INVALID, "width * height must be a multiple of 64 for multi-view renders");
    if ((size_t)n_views * spp * P >= ((size_t)1 << 32) - 1) return fail(MRP_E_INVALID, "too many items in one batch");
    const void* packed = nullptr;
    int rc = prepare_volume(host, &p0, &packed);
    if (rc) return rc;
    std::vector<NNArgs> nn((size_t)n_views);
    std::vector<GatherArgs> gv((size_t)n_views);
    for (int k = 0; k < n_views; ++k) {
        const mrp_render_params& p = ps[k];
        if (p.width != p0.width || p.height != p0.height) return fail(MRP_E_INVALID, "views differ in image size");
        if (p.mlp_precision != p0.mlp_precision || p.g != p0.g || p.scatter_rate.x != p0.scatter_rate.x ||
            p.scatter_rate.y != p0.scatter_rate.y || p.scatter_rate.z != p0.scatter_rate.z)
            return fail(MRP_E_INVALID, "views differ in network settings (precision, g or albedo)");
        Volume* v = vols[k];
        if (v->device != host->device) return fail(MRP_E_INVALID, "views on different devices");
        if (v->hdri) return fail(MRP_E_INVALID, "multi-view renders do not take environment-map lights");
        for (int j = 0; j < k; ++j)
            if (vols[j] == v && (ps[j].light_dir.x != p.light_dir.x || ps[j].light_dir.y != p.light_dir.y ||
                                 ps[j].light_dir.z != p.light_dir.z || ps[j].alpha != p.alpha || ps[j].seed != p.seed))
                return fail(MRP_E_INVALID, "views of one volume differ in light, alpha or seed");
        if (k > 0) {
            rc = prepare_tables(v, &p);
            if (rc) return rc;
        }
        image_args(v, &p, PixelMap{}, 0, &nn[k], &gv[k]);
    }
    // the tables were brought up to date on each volume's own stream
    for (int k = 1; k < n_views; ++k)
        if (vols[k]->stream != host->stream) MRP_HIP(hipStreamSynchronize(vols[k]->stream));
    NNArgs* nn_dev = (NNArgs*)tables_dev;
    GatherArgs* gv_dev = (GatherArgs*)(nn_dev + n_views);
    MRP_HIP(hipMemcpyAsync(nn_dev, nn.data(), nn.size() * sizeof(NNArgs), hipMemcpyHostToDevice, host->stream));
    MRP_HIP(hipMemcpyAsync(gv_dev, gv.data(), gv.size() * sizeof(GatherArgs), hipMemcpyHostToDevice, host->stream));
    NNArgs a = nn[0];
    a.counters = host->counters;
    a.views = nn_dev;
    a.view_items = (uint32_t)(spp * P);
    a.occ_words = 0u;
    for (int k = 0; k < n_views; ++k) a.occ_words = std::max(a.occ_words, (uint32_t)occupancy_words(nn[k].res));
    GatherArgs ga = gv[0];
    ga.views = gv_dev;
    ga.view_items = a.view_items;
    rc = run_stages(host, a, ga, p0.mlp_precision, packed, host->nn_weights, p0.g, srp_of(V(p0.scatter_rate)),
                    (size_t)n_views * spp * P, out, times, n_active, false);
    // the host copies of the tables must outlive the copies above
    MRP_HIP(hipStreamSynchronize(host->stream));
    return rc;
}

// Adds the pending chunk's descriptor count and stage times (the stream has
// been synchronised since that chunk was launched).
int mrpnn_finish(Volume* v, NNTimes* times, uint64_t* n_active) {
    if (!v->nn_pending) return MRP_OK;
    v->nn_pending = false;
    if (n_active) *n_active += (uint32_t)v->host_words[0];
    if (times) {
        const hipEvent_t t0 = v->nn_events[0], tc = v->nn_events[1], ts0 = v->nn_events[2], ts1 = v->nn_events[3],
                         t2 = v->nn_events[4], t3 = v->nn_events[5];
        float ms;
        MRP_HIP(hipEventElapsedTime(&ms, t0, tc));
        times->first_vertex += ms;
        MRP_HIP(hipEventElapsedTime(&ms, tc, ts0));
        times->compact += ms;
        MRP_HIP(hipEventElapsedTime(&ms, ts0, ts1));
        times->shadow += ms;
        MRP_HIP(hipEventElapsedTime(&ms, ts1, t2));
        times->gather += ms;
        MRP_HIP(hipEventElapsedTime(&ms, t2, t3));
        times->network += ms;
    }
    return MRP_OK;
}

int mrpnn_dev(Volume* v, const mrp_render_params* p, float* target, double* sum64, mrp_trace_stats* st) {
    const int W = p->width, H = p->height;
    const size_t P = (size_t)W * H;
    // with an environment map each sample has its own light (and transmittance
    // mips): one sample per batch
    const bool env_lights = v->hdri != nullptr;
    const int batch = env_lights ? 1
                                 : (int)std::max<size_t>(1, std::min<size_t>(v->sample_budget / (P * 16),
                                                                             (size_t)p->spp_count));
    int rc = ensure_samples(v, (size_t)batch * P * 16);
    if (rc) return rc;
    for (hipEvent_t& e : v->nn_events)
        if (!e) MRP_HIP(hipEventCreate(&e));
    const hipEvent_t e0 = v->nn_events[7], e1 = v->nn_events[8];
    MRP_HIP(hipEventRecord(e0, v->stream));
    rc = clear_counters(v);
    NNTimes tm;
    uint64_t n_active = 0;
    const PixelMap map;
    for (int s0 = 0; s0 < p->spp_count && rc == MRP_OK; s0 += batch) {
        const int n = std::min(batch, p->spp_count - s0);
        const bool last = s0 + n >= p->spp_count;
        mrp_render_params q = *p;
        if (env_lights) env_light_of_sample(v, p->spp_begin + s0, p->seed, &q.light_dir, &q.light_color);
        if (n == 1) {  // the combine stage folds the one sample into the running mean
            SampleFold fold;
            fold.target = target;
            fold.sum64 = sum64;
            fold.rate = 1.0f / (float)(1 + p->spp_begin + s0);
            rc = mrpnn_samples(v, &q, map, p->spp_begin + s0, 1, &tm, &n_active, last, fold);
        } else {
            rc = mrpnn_samples(v, &q, map, p->spp_begin + s0, n, &tm, &n_active, last);
            if (rc == MRP_OK) rc = launch_accumulate(v, v->samples, P, p->spp_begin + s0, n, target, sum64);
        }
    }
    float total_ms = 0.0f;
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    if (rc == MRP_OK) {
        // one host synchronisation per call: counters into pinned memory, then the end event
        MRP_HIP(hipMemcpyAsync(v->host_words + 1, v->counters, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               v->stream));
        MRP_HIP(hipEventRecord(e1, v->stream));
        MRP_HIP(hipEventSynchronize(e1));
        MRP_HIP(hipEventElapsedTime(&total_ms, e0, e1));
        for (int q = 0; q < 4; ++q) cnt4[q] = v->host_words[1 + q];
        rc = mrpnn_finish(v, &tm, &n_active);
    } else {
        hipStreamSynchronize(v->stream);
        v->nn_pending = false;
    }
    if (rc) return rc;
    if (st) {
        st->lookups = cnt4[0] + cnt4[2];
        st->shadow_lookups = cnt4[2];
        st->paths = (uint64_t)p->spp_count * P;
        st->scatters = n_active;  // first vertices that scattered: the descriptors gathered and evaluated
        st->max_path_steps = 0;
        st->kernel_ms = tm.network;
        st->reduce_ms = tm.compact + tm.gather;
        st->total_ms = total_ms;
        st->launches = 0;
        st->grid_blocks = 0;
        st->first_vertex_ms = tm.first_vertex;
        st->shadow_ms = tm.shadow;
        st->compact_ms = tm.compact;
        st->gather_ms = tm.gather;
    }
    return MRP_OK;
}

int mrpnn_rays(Volume* v, int n, const float* ori, const float* dir, const mrp_render_params* p, float* out,
               mrp_trace_stats* st) {
    if (!ori || !dir) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0 || n > (1 << 28)) return fail(MRP_E_INVALID, "n must be in [0, 2^28]");
    MRP_HIP(hipSetDevice(v->device));
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() must run first");
    if (n == 0) return MRP_OK;
    // `NNPredict` draws one sample per call (core/render.cuh:469-523): sample 0 of ray i's stream
    mrp_render_params q = *p;
    q.width = n;
    q.height = 1;
    q.spp_begin = 0;
    q.spp_count = 1;
    int rc = check_params(v, &q);
    if (rc) return rc;
    rc = ensure_samples(v, (size_t)n * 16);
    if (rc) return rc;
    rc = clear_counters(v);
    if (rc) return rc;
    PixelMap map;
    map.ray_ori = ori;
    map.ray_dir = dir;
    NNTimes tm;
    uint64_t n_active = 0;
    rc = mrpnn_samples(v, &q, map, 0, 1, &tm, &n_active, false);
    if (rc) return rc;
    MRP_HIP(hipMemcpyAsync(out, v->samples, (size_t)n * 16, hipMemcpyDeviceToDevice, v->stream));
    unsigned long long cnt4[4] = {0, 0, 0, 0};
    rc = read_counters(v, cnt4);
    if (rc) return rc;
    if (st) {
        *st = mrp_trace_stats{};
        st->lookups = cnt4[0] + cnt4[2];
        st->shadow_lookups = cnt4[2];
        st->paths = (uint64_t)n;
        st->scatters = n_active;
        st->kernel_ms = tm.network;
        st->first_vertex_ms = tm.first_vertex;
        st->shadow_ms = tm.shadow;
        st->compact_ms = tm.compact;
        st->gather_ms = tm.gather;
    }
    return MRP_OK;
}

}  // namespace mrp

using namespace mrp;

struct mrp_network {
    float* w = nullptr;    // FP32 weights (HBM)
    void* packed = nullptr;      // BF16 MFMA operand packing (HBM)
    void* packed_fp8 = nullptr;  // FP8-E4M3 MFMA operand packing and per-layer scales (HBM)
    int device = 0;
};

extern "C" {

int mrp_network_create(const float* weights, size_t count, mrp_network** out) {
    if (!weights || !out) return fail(MRP_E_INVALID, "NULL argument");
    *out = nullptr;
    if (count != (size_t)MRP_W_TOTAL) return fail(MRP_E_INVALID, "weight count must be 49673 (NNWeight.cuh layout)");
    mrp_network* n = new (std::nothrow) mrp_network();
    if (!n) return fail(MRP_E_NOMEM, "host allocation failed");
    hipGetDevice(&n->device);
    if (hipMalloc(&n->w, count * sizeof(float)) != hipSuccess) {
        delete n;
        return fail(MRP_E_NOMEM, "weight allocation failed");
    }
    if (hipMemcpy(n->w, weights, count * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(n->w);
        delete n;
        return fail(MRP_E_HIP, "weight upload failed");
    }
    int rc = pack_mlp(weights, 0, &n->packed);
    if (rc == MRP_OK) rc = pack_mlp(weights, 1, &n->packed_fp8);
    if (rc) {
        hipFree(n->w);
        hipFree(n->packed);
        delete n;
        return rc;
    }
    *out = n;
    return MRP_OK;
}

int mrp_network_destroy(mrp_network* n) {
    if (!n) return MRP_OK;
    hipFree(n->w);
    hipFree(n->packed);
    hipFree(n->packed_fp8);
    delete n;
    return MRP_OK;
}

// Shared set-up of the gather entry points on caller-given points: checks,
// tap tables, transmittance mips and HG table, and the gather's arguments.
static int gather_setup(mrp_volume* vol, int n, const float* pos_act_dev, const float* frame_x_dev,
                        const float* frame_y_dev, const float* frame_z_dev, mrp_vec3 light_dir, float alpha, float g,
                        uint64_t seed, GatherArgs* ga) {
    if (!vol || !pos_act_dev || !frame_x_dev || !frame_y_dev || !frame_z_dev) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0) return fail(MRP_E_INVALID, "n must be >= 0");
    Volume* v = &vol->v;
    MRP_HIP(hipSetDevice(v->device));
    if (!v->updated) return fail(MRP_E_STATE, "mrp_volume_update() has not run");
    int rc = upload_tap_tables(v->device);
    if (rc) return rc;
    rc = ensure_tr(v, light_dir, alpha);
    if (rc) return rc;
    rc = ensure_hg_lut(v, g, seed);
    if (rc) return rc;
    for (int m = 0; m < kNumMips; ++m) ga->mips[m] = v->mip_cells[m];
    for (int m = 0; m < kNumTrMips; ++m) ga->trs[m] = v->tr_cells[m];
    ga->lut = v->lut_cells;
    ga->mips16 = v->mip_cells16;
    ga->trs16 = v->tr_cells16;
    ga->light = normalize(V(light_dir));
    ga->lx = normalize(ga->light);
    ga->id = inv3(ga->lx);
    ga->tf = v3(0.4999f, 0.4999f, 0.4999f) * ga->id;
    ga->alpha = alpha;
    ga->g = g;
    ga->n_items = (uint32_t)n;
    ga->n_dev = nullptr;
    return MRP_OK;
}

int mrp_mrpnn_features(mrp_volume* vol, int n, const float* pos_act_dev, const float* frame_x_dev,
                       const float* frame_y_dev, const float* frame_z_dev, mrp_vec3 light_dir, float alpha, float g,
                       uint64_t seed, float* feat_dev, float* gamma_dev) {
    if (!feat_dev || !gamma_dev) return fail(MRP_E_INVALID, "NULL argument");
    GatherArgs ga;
    int rc = gather_setup(vol, n, pos_act_dev, frame_x_dev, frame_y_dev, frame_z_dev, light_dir, alpha, g, seed, &ga);
    if (rc || n == 0) return rc;
    Volume* v = &vol->v;
    gather_kernel<false><<<(n + kGatherItems - 1) / kGatherItems, kTaps, 0, v->stream>>>(
        ga, (const float4*)pos_act_dev, (const float4*)frame_x_dev, (const float4*)frame_y_dev,
        (const float4*)frame_z_dev, nullptr, feat_dev, gamma_dev, nullptr);
    MRP_HIP(hipGetLastError());
    MRP_HIP(hipStreamSynchronize(v->stream));
    return MRP_OK;
}

int mrp_mrpnn_features_tiled(mrp_volume* vol, int n, const float* pos_act_dev, const float* frame_x_dev,
                             const float* frame_y_dev, const float* frame_z_dev, mrp_vec3 light_dir, float alpha,
                             float g, uint64_t seed, void* tiles_dev, float* pool_dev, float* gamma_dev) {
    if (!tiles_dev || !pool_dev || !gamma_dev) return fail(MRP_E_INVALID, "NULL argument");
    GatherArgs ga;
    int rc = gather_setup(vol, n, pos_act_dev, frame_x_dev, frame_y_dev, frame_z_dev, light_dir, alpha, g, seed, &ga);
    if (rc || n == 0) return rc;
    Volume* v = &vol->v;
    rc = launch_gather_tiled(v, ga, (unsigned)((n + kGatherItems - 1) / kGatherItems), (const float4*)pos_act_dev,
                             (const float4*)frame_x_dev, (const float4*)frame_y_dev, (const float4*)frame_z_dev,
                             nullptr, tiles_dev, gamma_dev, pool_dev);
    if (rc) return rc;
    MRP_HIP(hipStreamSynchronize(v->stream));
    return MRP_OK;
}

int mrp_volume_set_gather_cells(mrp_volume* vol, int bits) {
    if (!vol) return fail(MRP_E_INVALID, "NULL volume");
    if (bits != 16 && bits != 32) return fail(MRP_E_INVALID, "gather cells must be 32 (FP32) or 16 (FP16) bits");
    vol->v.gather_cells16 = bits == 16;
    return MRP_OK;
}

int mrp_mlp_forward_tiled(const mrp_network* net, int n, const void* tiles_dev, const float* pool_dev,
                          const float* gamma_dev, float g, mrp_vec3 srp, int precision, float* out_dev) {
    if (!net || !tiles_dev || !pool_dev || !gamma_dev || !out_dev) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0) return fail(MRP_E_INVALID, "n must be >= 0");
    if (precision != 1 && precision != 2) return fail(MRP_E_INVALID, "tiled descriptors: precision 1 (BF16) or 2 (FP8)");
    if (n == 0) return MRP_OK;
    MRP_HIP(hipSetDevice(net->device));
    const void* packed = precision == 2 ? net->packed_fp8 : net->packed;
    if (!packed) return fail(MRP_E_STATE, "network has no MFMA packing");
    const int rc = launch_mlp(packed, precision - 1, nullptr, tiles_dev, pool_dev, gamma_dev, g, V(srp), (uint32_t)n,
                              nullptr, out_dev, nullptr);
    if (rc) return rc;
    MRP_HIP(hipDeviceSynchronize());
    return MRP_OK;
}

int mrp_mlp_forward(const mrp_network* net, int n, const float* feat_dev, const float* gamma_dev, float g,
                    mrp_vec3 srp, int precision, float* out_dev) {
    if (!net || !feat_dev || !gamma_dev || !out_dev) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0) return fail(MRP_E_INVALID, "n must be >= 0");
    if (n == 0) return MRP_OK;
    MRP_HIP(hipSetDevice(net->device));
    const vec3 s = V(srp);
    if (precision == 0) {
        mlp_fp32_kernel<<<(n + 63) / 64, 64>>>(net->w, feat_dev, gamma_dev, g, s, (uint32_t)n, nullptr, out_dev);
        MRP_HIP(hipGetLastError());
    } else if (precision == 1 || precision == 2) {
        const void* packed = precision == 2 ? net->packed_fp8 : net->packed;
        if (!packed) return fail(MRP_E_STATE, "network has no MFMA packing");
        const int rc =
            launch_mlp(packed, precision - 1, feat_dev, nullptr, nullptr, gamma_dev, g, s, (uint32_t)n, nullptr, out_dev,
                       nullptr);
        if (rc) return rc;
    } else {
        return fail(MRP_E_INVALID, "precision must be 0 (FP32), 1 (BF16) or 2 (FP8)");
    }
    MRP_HIP(hipDeviceSynchronize());
    return MRP_OK;
}

// Allocation-free network entry point.  The pooled squeeze-excitation
// statistics of the descriptors (72 floats each), which mrp_mlp_forward puts
// in a buffer it allocates per call, go to this fixed device array instead.
// Every GPU the library runs on holds its own copy (18 MiB) until the process
// ends; calls on one GPU share it, so calls that overlap on one GPU race.
// Every entry a call reads is written by the same call first, so nothing is
// cleared between calls.
constexpr uint32_t kNoAllocMaxItems = 65536;
static __device__ __align__(256) float scratch[(size_t)kNoAllocMaxItems * 72];
static float* g_scratch_dev[64] = {};
static std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address).
static int scratch_address(int device, float** out) {
    if (device < 0 || device >= 64) return fail(MRP_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(MRP_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_dev[device] = (float*)ptr;
    }
    *out = g_scratch_dev[device];
    return MRP_OK;
}

int mrp_mlp_forward_noalloc(const mrp_network* net, int n, const float* feat_dev, const float* gamma_dev, float g,
                            mrp_vec3 srp, int precision, float* out_dev) {
    if (!net || !feat_dev || !gamma_dev || !out_dev) return fail(MRP_E_INVALID, "NULL argument");
    if (n < 0) return fail(MRP_E_INVALID, "n must be >= 0");
    if ((uint32_t)n > kNoAllocMaxItems)
        return fail(MRP_E_INVALID, "n exceeds 65536, the descriptors the allocation-free scratch holds");
    if (precision < 0 || precision > 2) return fail(MRP_E_INVALID, "precision must be 0 (FP32), 1 (BF16) or 2 (FP8)");
    if (n == 0) return MRP_OK;
    MRP_HIP(hipSetDevice(net->device));
    const vec3 s = V(srp);
    if (precision == 0) {
        mlp_fp32_kernel<<<(n + 63) / 64, 64>>>(net->w, feat_dev, gamma_dev, g, s, (uint32_t)n, nullptr, out_dev);
        MRP_HIP(hipGetLastError());
    } else {
        const void* packed = precision == 2 ? net->packed_fp8 : net->packed;
        if (!packed) return fail(MRP_E_STATE, "network has no MFMA packing");
        float* pool = nullptr;
        int rc = scratch_address(net->device, &pool);
        if (rc) return rc;
        rc = launch_pool(feat_dev, (uint32_t)n, pool, nullptr);
        if (rc) return rc;
        rc = launch_mlp(packed, precision - 1, feat_dev, nullptr, pool, gamma_dev, g, s, (uint32_t)n, nullptr, out_dev,
                        nullptr);
        if (rc) return rc;
    }
    MRP_HIP(hipDeviceSynchronize());
    return MRP_OK;
}

int mrp_volume_set_network(mrp_volume* vol, const mrp_network* n) {
    if (!vol) return fail(MRP_E_INVALID, "NULL volume");
    vol->v.nn_weights = n ? n->w : nullptr;
    vol->v.nn_packed = n ? n->packed : nullptr;
    vol->v.nn_packed_fp8 = n ? n->packed_fp8 : nullptr;
    return MRP_OK;
}

}  // extern "C"
