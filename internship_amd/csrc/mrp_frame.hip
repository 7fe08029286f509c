// mrp_frame.hip - the progressive frame of `VolumeRender::Render(float3* target,
// Histogram* histo, unsigned* target2, ...)` (core/volume.cu:469-547) on gfx950.
//
// One call = one frame:
//   1. the path tracer (PT) or the MRPNN stages render the pixels of this call
//      (all, or one checkerboard half) into the per-pixel sample buffer, with
//      the progressive camera (row 0 at the bottom);
//   2. frame_update_kernel folds them into the caller's buffers: PT keeps the
//      running mean and the per-pixel histogram of r/(r+1) (`RenderCamera`,
//      core/volume.cu:99-123); MRPNN blends with the previous frame, fetched
//      through the reprojection into the previous camera (:124-143);
//   3. PT: denoise_pack_kernel - frame 0 fills the unrendered checkerboard
//      half from its 4 neighbours; later frames filter noisy pixels
//      (|E[l^2]-E[l]^2| > 0.01) with weights max(0, 1-1.2 chi2) over an 11x11
//      window of histograms (`Denoise`, :209-264), staged through LDS in
//      16x16 tiles with a 5-pixel apron;
//      MRPNN: reproject_pack_kernel stores the frame in the histogram buffer
//      for the next reprojection (`ReprojectionDenoise`, :266-287);
//   4. exposure, tone curve and the BGRA8 pack into the caller's buffer.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mrp_frame.h"
#include "mrp_internal.h"

namespace mrp {

struct FrameArgs {
    int width, height;
    int checker;        // parity rendered by this call, -1 without the checkerboard
    int checkerboard;   // the checkerboard setting
    int frame;          // the caller's frame number (`randseed`)
    int predict;        // MRPNN branch
    float exposure;
    int tone, denoise;
    // reprojection
    vec3 ori, forward, right, up;
    vec3 lori, lup, lright;
    uint32_t seed_lo, seed_hi, sample;
};

__global__ __launch_bounds__(256) void frame_update_kernel(FrameArgs a, const float4* __restrict__ samples,
                                                           float* __restrict__ target, float* __restrict__ histo) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= a.width * a.height) return;
    const int row = pix / a.width, col = pix - row * a.width;
    if (!checker_rendered(row, col, a.checker)) return;
    const float4 s = samples[pix];
    const vec3 res = vmax(v3(0, 0, 0), v3(s.x, s.y, s.z));
    const int fnum = frame_index(a.frame, a.checkerboard);
    float* t3 = target + 3 * (size_t)pix;
    if (!a.predict) {
        float h[kHistoFloats];
        float* hp = histo + (size_t)kHistoFloats * pix;
        for (int i = 0; i < kHistoFloats; ++i) h[i] = hp[i];
        float tg[3] = {t3[0], t3[1], t3[2]};
        histo_accumulate(tg, h, res, fnum);
        for (int i = 0; i < kHistoFloats; ++i) hp[i] = h[i];
        t3[0] = tg[0];
        t3[1] = tg[1];
        t3[2] = tg[2];
    } else {
        // the camera ray of this pixel (same jitter draws as the render)
        Rng rng = rng_make((uint32_t)pix, a.sample, 0u, ((uint64_t)a.seed_hi << 32) | a.seed_lo);
        const float xu = rng_next(&rng);
        const float xv = rng_next(&rng);
        const vec3 dir = camera_dir_rows(a.forward, a.right, a.up, row, col, a.width, a.height, xu, xv, 1);
        const bool sky = s.w < 0.0f;
        const float dis = fmax_(0.001f, sky ? 10.0f : s.w);
        const int lidx = reproject_pixel(a.ori + dir * dis, a.lori, a.lup, a.lright, a.width, a.height);
        float rate = 1.0f / (float)(1 + fnum);
        if (!sky && !a.checkerboard) rate = fmin_(0.2f, rate);
        const float* hl = histo + (size_t)kHistoFloats * lidx;
        const vec3 his = v3(hl[3 * kHistoBins], hl[3 * kHistoBins + 2], hl[3 * kHistoBins + 1]);
        const vec3 t = lerp(his, res, rate);
        t3[0] = t.x;
        t3[1] = t.y;
        t3[2] = t.z;
    }
}

constexpr int kTile = 16, kApron = 5, kSpan = kTile + 2 * kApron, kSpanCells = kSpan * kSpan;

// blockIdx.z selects one of several images of the same size stored one after
// another (batched serving); 0 for a single frame.
// BinT: the LDS type of the staged bin counts.  They are integer sample
// counts, so FP16 holds them exactly up to 2048: the host picks it when no
// count can exceed that (40 KB of bins instead of 81 KB: 3 workgroups per CU
// instead of 1), and FP32 otherwise.
template <typename BinT>
__global__ __launch_bounds__(256) void denoise_pack_kernel(FrameArgs a, const float* __restrict__ target,
                                                           const float* __restrict__ histo,
                                                           uint32_t* __restrict__ out) {
    __shared__ BinT s_bins[3 * kHistoBins][kSpanCells];  // bin-major: lanes read consecutive cells
    __shared__ float s_total[kSpanCells];
    __shared__ float s_tgt[3][kSpanCells];
    const int W = a.width, H = a.height;
    const size_t image = (size_t)blockIdx.z * (size_t)W * (size_t)H;
    target += 3 * image;
    histo += (size_t)kHistoFloats * image;
    out += image;
    const int x0 = blockIdx.x * kTile - kApron, y0 = blockIdx.y * kTile - kApron;
    const int tid = threadIdx.y * kTile + threadIdx.x;
    // the histograms are staged only if a pixel of this tile is filtered
    // (|E[l^2]-E[l]^2| > 0.01): a quiet tile reads its 16x16 variances and packs
    bool filter = a.frame != 0 && a.denoise;
    if (filter) {
        const int px = blockIdx.x * kTile + threadIdx.x, py = blockIdx.y * kTile + threadIdx.y;
        bool noisy = false;
        if (px < W && py < H) {
            const float* hp = histo + (size_t)kHistoFloats * ((size_t)py * W + px);
            const float x = hp[3 * kHistoBins + 2], x2 = hp[3 * kHistoBins + 1];
            noisy = fabs_(x2 - x * x) > 0.01f;
        }
        filter = __syncthreads_or(noisy) != 0;
    }
    for (int c = tid; c < kSpanCells; c += kTile * kTile) {
        const int cy = c / kSpan, cx = c - cy * kSpan;
        const int gy = clampi_(y0 + cy, 0, H - 1), gx = clampi_(x0 + cx, 0, W - 1);
        const size_t gp = (size_t)gy * W + gx;
        s_tgt[0][c] = target[3 * gp];
        s_tgt[1][c] = target[3 * gp + 1];
        s_tgt[2][c] = target[3 * gp + 2];
        if (filter) {
            const float* hp = histo + (size_t)kHistoFloats * gp;
            for (int b = 0; b < 3 * kHistoBins; ++b) s_bins[b][c] = (BinT)hp[b];
            s_total[c] = hp[3 * kHistoBins];
        }
    }
    __syncthreads();
    const int col = blockIdx.x * kTile + threadIdx.x, row = blockIdx.y * kTile + threadIdx.y;
    if (col >= W || row >= H) return;
    const size_t pix = (size_t)row * W + col;
    const int cc = (threadIdx.y + kApron) * kSpan + threadIdx.x + kApron;  // this pixel's cell
    vec3 res = v3(s_tgt[0][cc], s_tgt[1][cc], s_tgt[2][cc]);
    if (a.frame == 0) {
        if (a.checkerboard && !checker_rendered(row, col, a.checker)) {
            const int up = cc - kSpan, dn = cc + kSpan;  // rows clamp through the apron's clamped copies
            const vec3 s0 = v3(s_tgt[0][up], s_tgt[1][up], s_tgt[2][up]);
            const vec3 s1 = v3(s_tgt[0][dn], s_tgt[1][dn], s_tgt[2][dn]);
            const vec3 s2 = v3(s_tgt[0][cc - 1], s_tgt[1][cc - 1], s_tgt[2][cc - 1]);
            const vec3 s3 = v3(s_tgt[0][cc + 1], s_tgt[1][cc + 1], s_tgt[2][cc + 1]);
            res = (((s0 + s1) + s2) + s3) / 4.0f;
        }
    } else if (filter) {
        const float* hp = histo + (size_t)kHistoFloats * pix;
        const float x = hp[3 * kHistoBins + 2], x2 = hp[3 * kHistoBins + 1];
        const float variance = fabs_(x2 - x * x);
        if (variance > 0.01f) {
            float cb[3 * kHistoBins];
            for (int b = 0; b < 3 * kHistoBins; ++b) cb[b] = (float)s_bins[b][cc];
            const float nc = s_total[cc];
            vec3 acc = v3(0, 0, 0);
            float ws = 0.0f;
            for (int i = -kApron; i <= kApron; ++i) {
                for (int j = -kApron; j <= kApron; ++j) {
                    const int nb = cc + i * kSpan + j;
                    const float syx = sqrtf(s_total[nb] / nc), sxy = sqrtf(nc / s_total[nb]);
                    float p = 0.0f, d = 0.0f;
                    for (int b = 0; b < 3 * kHistoBins; ++b) {
                        const float hx = cb[b], hy = (float)s_bins[b][nb];
                        if (hx != 0.0f || hy != 0.0f) {
                            p += 1.0f;
                            const float t = syx * hx - sxy * hy;
                            d += t * t / (hx + hy);
                        }
                    }
                    const float w = histo_weight(p == 0.0f ? 0.0f : d / p);
                    acc = acc + v3(s_tgt[0][nb], s_tgt[1][nb], s_tgt[2][nb]) * w;
                    ws += w;
                }
            }
            res = v3(acc.x / ws, acc.y / ws, acc.z / ws);
        }
    }
    out[pix] = tone_pack(res, a.exposure, a.tone);
}

__global__ __launch_bounds__(256) void reproject_pack_kernel(FrameArgs a, const float* __restrict__ target,
                                                             float* __restrict__ histo, uint32_t* __restrict__ out) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= a.width * a.height) return;
    const vec3 t = v3(target[3 * (size_t)pix], target[3 * (size_t)pix + 1], target[3 * (size_t)pix + 2]);
    float* hp = histo + (size_t)kHistoFloats * pix;
    hp[3 * kHistoBins] = t.x;
    hp[3 * kHistoBins + 2] = t.y;
    hp[3 * kHistoBins + 1] = t.z;
    out[pix] = tone_pack(t, a.exposure, a.tone);
}

__global__ void clear_histo_kernel(float* __restrict__ histo, size_t n_floats) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_floats) histo[i] = 0.0f;
}

// The denoise and pack of n_images images of one size (a batch of views):
// `Denoise<denoise>` past frame 0 on the running means `target` and the
// histograms `histo`, then exposure, tone curve and BGRA8.
int launch_denoise_images(hipStream_t stream, int width, int height, int n_images, int denoise, int tone,
                          float exposure, const float* target, const float* histo, uint32_t* bgra, int max_count) {
    FrameArgs a{};
    a.width = width;
    a.height = height;
    a.checker = -1;
    a.checkerboard = 0;
    a.frame = 1;
    a.predict = 0;
    a.exposure = exposure;
    a.tone = tone;
    a.denoise = denoise;
    const dim3 grid((unsigned)((width + kTile - 1) / kTile), (unsigned)((height + kTile - 1) / kTile),
                    (unsigned)n_images);
    if (max_count <= 2048)
        denoise_pack_kernel<_Float16><<<grid, dim3(kTile, kTile), 0, stream>>>(a, target, histo, bgra);
    else
        denoise_pack_kernel<float><<<grid, dim3(kTile, kTile), 0, stream>>>(a, target, histo, bgra);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

static vec3 V(const mrp_vec3& a) { return v3(a.x, a.y, a.z); }

static int render_frame(Volume* v, const mrp_frame_params* fp, float* target, float* histo, uint32_t* bgra,
                        mrp_trace_stats* st) {
    const mrp_render_params* p = &fp->r;
    int rc = check_params(v, p);
    if (rc) return rc;
    if (p->spp_count != 1) return fail(MRP_E_INVALID, "a progressive frame renders one sample (spp_count = 1)");
    if (fp->tone < 0 || fp->tone > 2) return fail(MRP_E_INVALID, "tone must be 0 (none), 1 (gamma) or 2 (ACES)");
    const int W = p->width, H = p->height;
    const size_t P = (size_t)W * H;
    rc = ensure_samples(v, P * sizeof(float4));
    if (rc) return rc;
    const bool predict = p->render_type == MRP_RENDER_MRPNN;
    PixelMap map;
    map.checker = fp->checkerboard ? v->frame_flip : -1;
    map.rows_up = 1;

    FrameArgs a;
    a.width = W;
    a.height = H;
    a.checker = map.checker;
    a.checkerboard = fp->checkerboard ? 1 : 0;
    a.frame = p->spp_begin;
    a.predict = predict ? 1 : 0;
    a.exposure = fp->exposure;
    a.tone = fp->tone;
    a.denoise = fp->denoise ? 1 : 0;
    a.ori = V(p->ori);
    a.forward = normalize(-a.ori);
    a.right = V(p->right);
    a.up = V(p->up);
    // before any previous frame there is nothing to reproject from: use this camera
    a.lori = v->frame_have_last ? V(v->frame_last_ori) : a.ori;
    a.lup = v->frame_have_last ? V(v->frame_last_up) : a.up;
    a.lright = v->frame_have_last ? V(v->frame_last_right) : a.right;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.sample = v->frame_calls;

    hipEvent_t e0, e1, e2;
    MRP_HIP(hipEventCreate(&e0));
    MRP_HIP(hipEventCreate(&e1));
    MRP_HIP(hipEventCreate(&e2));
    MRP_HIP(hipEventRecord(e0, v->stream));
    rc = clear_counters(v);
    NNTimes tm;
    uint64_t n_active = 0;
    if (rc == MRP_OK) {
        if (predict && !v->frame_last_predict) {
            const size_t nf = P * kHistoFloats;
            clear_histo_kernel<<<(unsigned)((nf + 255) / 256), 256, 0, v->stream>>>(histo, nf);
            MRP_HIP(hipGetLastError());
        }
        mrp_render_params q = *p;  // MRPNN with an environment map: this call's light
        if (predict) env_light_of_sample(v, (int)v->frame_calls, p->seed, &q.light_dir, &q.light_color);
        rc = predict ? mrpnn_samples(v, &q, map, (int)v->frame_calls, 1, &tm, &n_active)
                     : trace_samples(v, p, map, (int)v->frame_calls, 1);
    }
    if (rc == MRP_OK) {
        MRP_HIP(hipEventRecord(e1, v->stream));
        frame_update_kernel<<<(unsigned)((P + 255) / 256), 256, 0, v->stream>>>(a, (const float4*)v->samples, target,
                                                                             histo);
        MRP_HIP(hipGetLastError());
        if (!predict) {
            const dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);
            // a pixel's counts are at most its frames since its histogram was
            // cleared (frame index 0), and at most the volume's calls since its
            // last reset (a caller that never passes frame 0)
            const uint32_t fnum = (uint32_t)frame_index(a.frame, a.checkerboard);
            if (fnum < 2048u && v->frame_calls < 2048u)
                denoise_pack_kernel<_Float16><<<grid, dim3(kTile, kTile), 0, v->stream>>>(a, target, histo, bgra);
            else
                denoise_pack_kernel<float><<<grid, dim3(kTile, kTile), 0, v->stream>>>(a, target, histo, bgra);
        } else {
            reproject_pack_kernel<<<(unsigned)((P + 255) / 256), 256, 0, v->stream>>>(a, target, histo, bgra);
        }
        MRP_HIP(hipGetLastError());
        MRP_HIP(hipEventRecord(e2, v->stream));
        MRP_HIP(hipEventSynchronize(e2));
    }
    float render_ms = 0.0f, post_ms = 0.0f;
    unsigned long long cnt[4] = {0, 0, 0, 0};
    if (rc == MRP_OK) {
        MRP_HIP(hipEventElapsedTime(&render_ms, e0, e1));
        MRP_HIP(hipEventElapsedTime(&post_ms, e1, e2));
        rc = read_counters(v, cnt);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (rc) return rc;
    // state for the next call (core/volume.cu:506-507,517-519,537-545)
    v->frame_flip = (v->frame_flip + 1) % 2;
    v->frame_calls += 1;
    v->frame_last_predict = predict;
    v->frame_have_last = true;
    v->frame_last_ori = p->ori;
    v->frame_last_up = p->up;
    v->frame_last_right = p->right;
    if (st) {
        st->lookups = predict ? cnt[0] + cnt[2] : cnt[0];  // MRPNN: shadow-ray lookups are counted apart
        st->shadow_lookups = predict ? cnt[2] : 0;
        st->paths = predict ? (uint64_t)(fp->checkerboard ? (P + 1) / 2 : P) : cnt[1];
        st->scatters = predict ? n_active : cnt[2];
        st->max_path_steps = cnt[3];
        st->kernel_ms = predict ? tm.network : (double)render_ms;
        st->reduce_ms = post_ms;  // histogram update + denoise + tone + pack
        st->total_ms = (double)render_ms + (double)post_ms;
        st->first_vertex_ms = tm.first_vertex;
        st->shadow_ms = tm.shadow;
        st->compact_ms = tm.compact;
        st->gather_ms = tm.gather;
        st->launches = 1;
        st->grid_blocks = 0;
    }
    return MRP_OK;
}

}  // namespace mrp

using namespace mrp;

extern "C" {

int mrp_render_frame(mrp_volume* vol, const mrp_frame_params* p, float* target_dev, mrp_histogram* histo_dev,
                     uint32_t* bgra_dev, mrp_trace_stats* stats) {
    if (!vol || !p || !target_dev || !histo_dev || !bgra_dev) return fail(MRP_E_INVALID, "NULL argument");
    MRP_HIP(hipSetDevice(vol->v.device));
    return render_frame(&vol->v, p, target_dev, (float*)histo_dev, bgra_dev, stats);
}

int mrp_volume_reset_frames(mrp_volume* vol) {
    if (!vol) return fail(MRP_E_INVALID, "NULL argument");
    Volume* v = &vol->v;
    v->frame_flip = 0;
    v->frame_calls = 0;
    v->frame_have_last = false;
    v->frame_last_predict = true;
    return MRP_OK;
}

int mrp_volume_frame_state(const mrp_volume* vol, int* flip, uint32_t* calls) {
    if (!vol || !flip || !calls) return fail(MRP_E_INVALID, "NULL argument");
    *flip = vol->v.frame_flip;
    *calls = vol->v.frame_calls;
    return MRP_OK;
}

}  // extern "C"
