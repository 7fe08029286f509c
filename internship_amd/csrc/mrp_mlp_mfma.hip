// mrp_mlp_mfma.hip - the MRPNN network on gfx950 matrix cores: BF16 or FP8-E4M3
// operands, FP32 accumulation.
//
// The reference evaluates the network (`RadiancePredict` lines 370-486,
// core/radiancePredict.cu) one pixel per thread with scalar loops.  Here one
// wave evaluates 32 descriptors at once: every layer is a v_mfma_f32_32x32x16_bf16
// product D[32 out x 32 items] = W[32 x 16] x X[16 x 32 items], tiled over the
// layer's rows (32) and inputs (16).
//
// Register layout ("D-order").  Lane l works on item n = l % 32 and lane group
// kg = l / 32.  An MFMA result leaves row r = 8*(v/4) + 4*kg + (v%4) of item n in
// register v (v = 0..15).  The next layer consumes those registers directly as
// its B operand (K block b = registers 8b..8b+7) because its weight columns are
// permuted on the host to the same order: A element e of K block kb = 2j+b for
// lane group kg multiplies input feature 32j + 16b + 8*(e/4) + 4*kg + (e%4).
// Descriptor features and pooled statistics are loaded from HBM straight into
// that order (two float4 per 16-feature block), so no layer needs a shuffle.
// Layers narrower than 32x16 are zero-padded (rows >= M and features >= K have
// zero weights; padded inputs are kept finite).
//
// Squeeze-excitation gates (8x8 -> ReLU -> 3x8 -> sigmoid) also run as MFMA
// products; the three gates are exchanged between the two lane groups of an
// item with one cross-lane move.  The per-segment means and maxima the gates
// start from come precomputed from the gather kernel.
//
// FP8 (v_mfma_f32_32x32x16_fp8_fp8, OCP E4M3) uses the same tiles and register
// order with 8 one-byte elements per lane.  Scales are powers of two, so scaling
// and unscaling are exact:
//   - weights: one scale per layer, 2^(7 - floor(log2 max|W|)), which puts the
//     layer's largest weight in [128, 256) (E4M3 reaches 448);
//   - activations: one scale per layer and descriptor, chosen the same way
//     from the descriptor's largest input magnitude to that layer (both lane
//     groups of the item, one cross-lane max).  Each lane's accumulator
//     column belongs to one descriptor, so unscaling is one multiply per
//     output: D = (W's x X's) * 2^-(sw + sx) + bias.
// Weights and activations are rounded by the same hardware conversion
// (v_cvt_pk_fp8_f32, round to nearest even), so both operands share one
// byte format by construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mrp_internal.h"
#include "mrp_mfma.h"
#include "mrp_mlp.h"
#include "mrp_mrpnn.h"

namespace mrp {

constexpr LayerTable kLT = make_layer_table();  // tile and bias offsets, as constants
static constexpr int bias_tiles_of(const LayerTable& t) {
    int b = 0;
    (void)total_tiles(t.L, &b);
    return b;
}
static constexpr int tiles_of(const LayerTable& t) {
    int b = 0;
    return total_tiles(t.L, &b);
}
static_assert(tiles_of(kLT) == kTiles && bias_tiles_of(kLT) == kBiasTiles,
              "the layer table's tiles must fill the kernels' LDS layout exactly");

// Pre-scaled FP32 weight tiles -> E4M3 tiles, with the instruction that also
// converts the activations.
__global__ void fp8_tiles_kernel(const float* __restrict__ src, long* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* s = src + (size_t)i * 8;
    dst[i] = pack8_fp8(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]);
}

int pack_mlp(const float* w, int fmt, void** packed_dev) {
    if (fmt != (int)Fmt::BF16 && fmt != (int)Fmt::FP8) return fail(MRP_E_INVALID, "unknown packed network format");
    const bool fp8 = fmt == (int)Fmt::FP8;
    LayerDesc L[LY_COUNT];
    layer_table(L);
    int nb = 0;
    const int nt = total_tiles(L, &nb);
    if (nt != kTiles || nb != kBiasTiles) return fail(MRP_E_STATE, "packed weight tiles do not match the kernel's LDS layout");
    std::vector<float> a((size_t)nt * 64 * 8, 0.0f);  // FP32 tiles (FP8: already scaled)
    std::vector<float> bias((size_t)nb * 32, 0.0f);
    float unscale[kScaleFloats];
    for (int i = 0; i < kScaleFloats; ++i) unscale[i] = 1.0f;
    for (int i = 0; i < LY_COUNT; ++i) {
        const LayerDesc& d = L[i];
        float sw = 1.0f;
        if (fp8) {
            float m = 0.0f;
            for (int r = 0; r < d.M * d.K; ++r) m = std::max(m, std::fabs(w[d.w + r]));
            if (m > 0.0f) {
                int e = 0;
                (void)std::frexp(m, &e);  // m = f * 2^e, f in [0.5, 1): m * 2^(8-e) lies in [128, 256)
                sw = std::ldexp(1.0f, 8 - e);
                unscale[i] = std::ldexp(1.0f, e - 8);
            }
        }
        const int mtn = (d.M + 31) / 32, kbn = (d.K + 15) / 16;
        for (int mt = 0; mt < mtn; ++mt)
            for (int kb = 0; kb < kbn; ++kb) {
                float* t = &a[((size_t)(d.tile + mt * kbn + kb)) * 512];
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int m = mt * 32 + l % 32, f = korder(kb, l / 32, e);
                        t[l * 8 + e] = (m < d.M && f < d.K) ? w[d.w + m * d.K + f] * sw : 0.0f;
                    }
            }
        if (d.b >= 0)
            for (int mt = 0; mt < mtn; ++mt)
                for (int kg = 0; kg < 2; ++kg)
                    for (int v = 0; v < 16; ++v) {
                        const int r = mt * 32 + vrow(v, kg);
                        bias[((size_t)(d.btile + mt) * 2 + kg) * 16 + v] = r < d.M ? w[d.b + r] : 0.0f;
                    }
    }
    float extra[kExtraFloats] = {};
    for (int i = 0; i < 24; ++i) extra[i] = w[MRP_W_LGGSW + i];
    for (int i = 0; i < 8; ++i) extra[24 + i] = w[MRP_W_LGGSB + i];
    for (int i = 0; i < 16; ++i) extra[32 + i] = w[MRP_W_LX12W + i];
    extra[48] = w[MRP_W_LX12B];
    const size_t abytes = fp8 ? tile_bytes<Fmt::FP8>() : tile_bytes<Fmt::BF16>(), bbytes = bias.size() * 4;
    void* dev = nullptr;
    if (hipMalloc(&dev, abytes + bbytes + sizeof(extra) + sizeof(unscale)) != hipSuccess)
        return fail(MRP_E_NOMEM, "packed weight allocation failed");
    hipError_t e = hipSuccess;
    if (fp8) {
        float* tmp = nullptr;
        e = hipMalloc(&tmp, a.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(tmp, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            const int n = nt * 64;
            fp8_tiles_kernel<<<(n + 255) / 256, 256>>>(tmp, (long*)dev, n);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
        if (tmp) hipFree(tmp);
    } else {
        std::vector<uint16_t> h(a.size());
        for (size_t i = 0; i < a.size(); ++i) h[i] = to_bf16_rne(a[i]);
        e = hipMemcpy(dev, h.data(), abytes, hipMemcpyHostToDevice);
    }
    char* p = (char*)dev + abytes;
    if (e == hipSuccess) e = hipMemcpy(p, bias.data(), bbytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + bbytes, extra, sizeof(extra), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + bbytes + sizeof(extra), unscale, sizeof(unscale), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(dev);
        return hip_fail(e, "packed weight upload");
    }
    *packed_dev = dev;
    return MRP_OK;
}

template <Fmt F, int MT, int KB>
__device__ __forceinline__ void dense(const Packed<F>& pk, int layer, const f32x16* X, f32x16 (&D)[MT], bool relu,
                                      int lane) {
    dense_d<F, MT, KB>(pk, kLT.L[layer], layer, X, D, relu, lane);
}

// The squeeze-excitation input of block `blk`: t = [mean A, mean B, mean C,
// max A, max B, max C, g, gamma]; lane group kg holds t[4kg..4kg+3].
__device__ __forceinline__ f32x16 se_input(int blk, const float* poolrow, float g, float gamma, int kg) {
    f32x16 t = zero16();
    if (kg == 0) {
        t[0] = poolrow[3 * blk];
        t[1] = poolrow[3 * blk + 1];
        t[2] = poolrow[3 * blk + 2];
        t[3] = poolrow[36 + 3 * blk];
    } else {
        t[0] = poolrow[36 + 3 * blk + 1];
        t[1] = poolrow[36 + 3 * blk + 2];
        t[2] = g;
        t[3] = gamma;
    }
    return t;
}

// The sigmoid gates of the BF16 / FP8 networks: 1 / (1 + 2^(-x log2 e)) on the
// hardware exponential and reciprocal (about 1 ulp each, against the 2^-9
// rounding of the BF16 products that the gates scale).
__device__ __forceinline__ float gate_sigmoid(float x) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * x));
}

// Squeeze-excitation gates of block `blk` from its input t: returns the 3
// sigmoid gates of this lane's item.
template <Fmt F>
__device__ __forceinline__ void se_gates(const Packed<F>& pk, int blk, f32x16 t, int lane, float* gate) {
    f32x16 h[1];
    dense<F, 1, 1>(pk, LY_SE0 + blk, &t, h, true, lane);
    f32x16 o[1];
    dense<F, 1, 1>(pk, LY_SE1 + blk, h, o, false, lane);
    // rows 0..2 live in lane group 0 (registers 0..2)
    const int src = lane & 31;
    gate[0] = gate_sigmoid(__shfl(o[0][0], src));
    gate[1] = gate_sigmoid(__shfl(o[0][1], src));
    gate[2] = gate_sigmoid(__shfl(o[0][2], src));
}

// ---------------------------------------------------------------- kernel
// Persistent: one workgroup of 8 waves per CU stages all weight tiles (BF16
// 150 KB, FP8 75 KB), the bias tiles and the small VALU layers into LDS once,
// then each wave evaluates tiles of 32 descriptors, reading every A operand
// from LDS.

struct MlpArgs {
    const uint4* packed;
    const float* feat;    // FP32 rows, or
    const uint4* tiled;   // wave-tiled BF16 descriptors (tile_chunk)
    const float* pool;
    const float* gamma;
    float g;
    vec3 srp;
    uint32_t n_items;
    const uint32_t* n_dev;  // if set, the descriptor count is read here (n_items bounds the grid)
    float* pred;
};


// The three streams' chunks of S features from feature `from` of a
// wave-tiled descriptor (lane group kg's half): cq the first 16 (8 if S is 8),
// cr the next 16 when S is 32.
__device__ __forceinline__ void stream_chunks(const uint4* __restrict__ tiled, uint32_t item, int from, int S, int kg,
                                              uint4 (&cq)[3], uint4 (&cr)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint4* tc = tiled + tile_chunk(item, k, from >> 4, kg);
        if (S == 8) {
            const uint2 h = ((const uint2*)tc)[(from >> 3) & 1];
            cq[k] = make_uint4(h.x, h.y, 0u, 0u);
        } else {
            cq[k] = tc[0];
        }
        cr[k] = S == 32 ? tc[64] : make_uint4(0u, 0u, 0u, 0u);  // the next chunk of the stream
    }
}

template <Fmt F, bool kTiled>
__device__ __forceinline__ void mlp_persistent(const MlpArgs& args, typename OpOf<F>::T* s_a, float* s_bias,
                                               float* s_extra, float* s_unscale) {
    {
        const uint4* packed = args.packed;
        constexpr int kA16 = (int)(tile_bytes<F>() / 16);
        uint4* da = (uint4*)s_a;
        stage16<kMlpBlock>(packed, da, kA16);
        const float4* sb = (const float4*)(packed + kA16);
        float4* db = (float4*)s_bias;
        for (int i = threadIdx.x; i < kBiasTiles * 8; i += kMlpBlock) db[i] = sb[i];
        const float* se = (const float*)(packed + kA16) + kBiasTiles * 32;
        for (int i = threadIdx.x; i < kExtraFloats; i += kMlpBlock) s_extra[i] = se[i];
        if (s_unscale)
            for (int i = threadIdx.x; i < kScaleFloats; i += kMlpBlock) s_unscale[i] = se[kExtraFloats + i];
    }
    __syncthreads();
    const float* __restrict__ feat = args.feat;
    const float* __restrict__ pool = args.pool;
    const float g = args.g;
    const uint32_t n_items = args.n_dev ? *args.n_dev : args.n_items;
    if (n_items == 0u) return;  // uniform; after the staging barrier
    const int lane = (int)__lane_id();
    const int kg = lane >> 5;
    const float* pextra = s_extra;
    const uint32_t n_tiles = (n_items + 31u) / 32u;
    for (uint32_t tile = blockIdx.x * kMlpWaves + (threadIdx.x >> 6); tile < n_tiles;
         tile += gridDim.x * kMlpWaves) {
    // The weight and bias tiles are the same for every tile of descriptors, and
    // their offsets are constants: an opaque zero offset, new each iteration,
    // keeps their LDS reads inside the loop (all of them read ahead would need
    // about 600 more registers).
    uint32_t opaque = 0u;
    asm volatile("" : "+s"(opaque));
    const Packed<F> pk{s_a + opaque, s_bias + opaque, s_unscale};
    const uint32_t item0 = tile * 32u;
    const uint32_t item = min(item0 + (uint32_t)(lane & 31), n_items - 1u);  // tail lanes recompute the last item
    const float* frow = kTiled ? nullptr : feat + (size_t)item * kFeatures;
    const float* prow = pool + (size_t)item * 72;
    const float gamma = args.gamma[item];

    // ------------------------------------------------ the two descriptor streams
    // prev[k]: previous block output of stream k in D-order (X_ValA2/B2/C2 window)
    f32x16 prev[3] = {zero16(), zero16(), zero16()};
    f32x16 comb[3];    // L81 outputs (32 each)
    float lout[3][4];  // LD41 outputs (8 rows: registers 0..3 of each lane group)
    const int from[12] = {0, 8, 16, 32, 48, 64, 96, 128, 160, 168, 176, 184};
    const int size[12] = {8, 8, 16, 16, 16, 32, 32, 32, 8, 8, 8, 8};
    // Loads run one block ahead: block blk + 1's gate input and stream chunks
    // are issued before block blk's products, so their latency overlaps a
    // whole block (loads complete in order; a wait for this block's values
    // leaves the next block's in flight).
    f32x16 t_next = se_input(0, prow, g, gamma, kg);
    uint4 cq_next[3] = {}, cr_next[3] = {};
    if constexpr (kTiled) stream_chunks(args.tiled, item, from[0], size[0], kg, cq_next, cr_next);
#pragma unroll
    for (int blk = 0; blk < 12; ++blk) {
        const int S = size[blk];
        const bool residual = (blk != 0 && blk != 8);
        const bool rep = (blk == 1 || blk == 4);  // REP3 after this block, ADD3 otherwise
        const f32x16 t_se = t_next;
        uint4 cq[3], cr[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cq[k] = cq_next[k];
            cr[k] = cr_next[k];
        }
        if (blk + 1 < 12) {
            t_next = se_input(blk + 1, prow, g, gamma, kg);
            if constexpr (kTiled)
                stream_chunks(args.tiled, item, from[blk + 1], size[blk + 1], kg, cq_next, cr_next);
        }
        float gate[3];
        se_gates<F>(pk, blk, t_se, lane, gate);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f32x16 xv = zero16();
            if constexpr (kTiled) {
                unpack_bf16(cq[k].x, xv, 0);
                unpack_bf16(cq[k].y, xv, 2);
                if (S != 8) {
                    unpack_bf16(cq[k].z, xv, 4);
                    unpack_bf16(cq[k].w, xv, 6);
                }
                if (S == 32) {
                    unpack_bf16(cr[k].x, xv, 8);
                    unpack_bf16(cr[k].y, xv, 10);
                    unpack_bf16(cr[k].z, xv, 12);
                    unpack_bf16(cr[k].w, xv, 14);
                }
            } else {
                const float* srow = frow + k * kTaps + from[blk];
                if (S == 8) {
                    const float4 q = *(const float4*)(srow + 4 * kg);
                    xv[0] = q.x;
                    xv[1] = q.y;
                    xv[2] = q.z;
                    xv[3] = q.w;
                } else {
                    load16(srow, 0, kg, xv, 0);
                    if (S == 32) load16(srow, 16, kg, xv, 1);
                }
            }
            // registers 0..R-1 hold the block's S input features (two lane groups,
            // 4 rows per register group); the others are padding, kept at 0 in the
            // BF16 form (FP8 computes every row: selecting the live rows took its
            // network from 0.21 to 0.25 ms at 256 VGPRs; the padding rows meet
            // zero weights either way)
            const int R = (S == 8 ? 4 : (S == 16 ? 8 : 16));
            constexpr bool kAllRows = F == Fmt::FP8;
            f32x16 x;
#pragma unroll
            for (int v = 0; v < 16; ++v)  // gate and residual in one rounding (one FMA)
                x[v] = (kAllRows || v < R) ? (residual ? fmaf(xv[v], gate[k], prev[k][v]) : xv[v] * gate[k]) : 0.0f;
            f32x16 y[1];
            if (S == 32) {
                dense<F, 1, 2>(pk, LY_STREAM + 3 * blk + k, &x, y, true, lane);
            } else {
                dense<F, 1, 1>(pk, LY_STREAM + 3 * blk + k, &x, y, true, lane);
            }
            if (rep) {
                // window = [block output | this block's input features]
                f32x16 w = y[0];
                if (S == 8) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) w[4 + v] = xv[v];
                } else {
#pragma unroll
                    for (int v = 0; v < 8; ++v) w[8 + v] = xv[v];
                }
                prev[k] = w;
            } else {
#pragma unroll
                for (int v = 0; v < 16; ++v) prev[k][v] = (kAllRows || v < R) ? y[0][v] + xv[v] : 0.0f;
            }
        }
        if (blk == 7) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                f32x16 o[1];
                dense<F, 1, 2>(pk, LY_L81 + k, &prev[k], o, true, lane);
                comb[k] = o[0];
                prev[k] = zero16();
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f32x16 o[1];
        dense<F, 1, 1>(pk, LY_LD41 + k, &prev[k], o, true, lane);
#pragma unroll
        for (int v = 0; v < 4; ++v) lout[k][v] = o[0][v];
    }

    // ------------------------------------------------ head, once per distinct channel
    const vec3 srp = args.srp;
    const float sr[3] = {srp.x, srp.y, srp.z};
    float P[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        if (c > 0 && sr[c] == sr[0]) { P[c] = P[0]; continue; }
        if (c == 2 && sr[2] == sr[1]) { P[2] = P[1]; continue; }
        // SE of the head: LSEFin1 (16x75) on the pool + [g, gamma, srp], LSEFin2 (6x16)
        f32x16 tp[3];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            load16(prow, 32 * j, kg, tp[j], 0);
            load16(prow, 32 * j + 16, kg, tp[j], 1);
        }
        {
            f32x16 t = zero16();
            // features 64..79: 64+8*(e/4)+4kg+e%4; 72 = g, 73 = gamma, 74 = srp_c
            if (kg == 0) {
                const float4 q = *(const float4*)(prow + 64);
                t[0] = q.x;
                t[1] = q.y;
                t[2] = q.z;
                t[3] = q.w;
                t[4] = g;
                t[5] = gamma;
                t[6] = sr[c];
            } else {
                const float4 q = *(const float4*)(prow + 68);
                t[0] = q.x;
                t[1] = q.y;
                t[2] = q.z;
                t[3] = q.w;
            }
            tp[2] = t;
        }
        f32x16 s1[1];
        dense<F, 1, 5>(pk, LY_FIN1, tp, s1, true, lane);
        f32x16 s2[1];
        dense<F, 1, 1>(pk, LY_FIN2, s1, s2, false, lane);
        float s[6];
        const int src0 = lane & 31, src1 = (lane & 31) + 32;
        s[0] = gate_sigmoid(__shfl(s2[0][0], src0));
        s[1] = gate_sigmoid(__shfl(s2[0][1], src0));
        s[2] = gate_sigmoid(__shfl(s2[0][2], src0));
        s[3] = gate_sigmoid(__shfl(s2[0][3], src0));
        s[4] = gate_sigmoid(__shfl(s2[0][0], src1));
        s[5] = gate_sigmoid(__shfl(s2[0][1], src1));
        // head input: [comb A * s0 | comb B * s1 | comb C * s2 | LD41 x s3..s5, LGGS]
        f32x16 x0[4];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int v = 0; v < 16; ++v) x0[j][v] = comb[j][v] * s[j];
        }
        // LGGS (8x3) rows 4kg..4kg+3 on [g, gamma, srp_c]
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 4 * kg + v;
            float acc = 0.0f;
            acc = fmaf(g, pextra[r * 3 + 0], acc);
            acc = fmaf(gamma, pextra[r * 3 + 1], acc);
            acc = fmaf(sr[c], pextra[r * 3 + 2], acc);
            acc += pextra[24 + r];
            x0[3][12 + v] = acc >= 0.0f ? acc : 0.0f;
            x0[3][v] = lout[0][v] * s[3];
            x0[3][4 + v] = lout[1][v] * s[4];
            x0[3][8 + v] = lout[2][v] * s[5];
        }
        f32x16 h0[4];
        dense<F, 4, 8>(pk, LY_LC0, x0, h0, true, lane);
        f32x16 h1[2];
        dense<F, 2, 8>(pk, LY_LC1, h0, h1, true, lane);
        f32x16 h2[1];
        dense<F, 1, 4>(pk, LY_LC2, h1, h2, true, lane);
        f32x16 a[1];
        dense<F, 1, 2>(pk, LY_LX, h2, a, true, lane);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            f32x16 hb[1];
            dense<F, 1, 1>(pk, LY_LXR + 2 * r, a, hb, true, lane);
            f32x16 hc[1];
            dense<F, 1, 1>(pk, LY_LXR + 2 * r + 1, hb, hc, true, lane);
#pragma unroll
            for (int v = 0; v < 8; ++v) a[0][v] = a[0][v] + hc[0][v];
        }
        // LX12 (1x16): rows 0..15 are registers 0..7 split over the two lane groups
        float part = 0.0f;
#pragma unroll
        for (int v = 0; v < 8; ++v) part = fmaf(a[0][v], pextra[32 + vrow(v, kg)], part);
        part += __shfl_xor(part, 32);
        const float pc = part + pextra[48];
        P[c] = pc >= 0.0f ? pc : 0.0f;
    }
    if (kg == 0 && item0 + (uint32_t)lane < n_items) {
        float* pred = args.pred;
        pred[3 * item] = fmax_(expf_(P[0]) - 1.0f, 0.0f) * sr[0];
        pred[3 * item + 1] = fmax_(expf_(P[1]) - 1.0f, 0.0f) * sr[1];
        pred[3 * item + 2] = fmax_(expf_(P[2]) - 1.0f, 0.0f) * sr[2];
    }
    }  // tile loop
}

// Renderer: wave-tiled BF16 descriptors.
__global__ __launch_bounds__(kMlpBlock) void mlp_bf16_kernel(MlpArgs args) {
    __shared__ bf16x8 s_a[kTiles * 64];
    __shared__ float s_bias[kBiasTiles * 32];
    __shared__ float s_extra[kExtraFloats];
    mlp_persistent<Fmt::BF16, true>(args, s_a, s_bias, s_extra, nullptr);
}

__global__ __launch_bounds__(kMlpBlock) void mlp_fp8_kernel(MlpArgs args) {
    __shared__ long s_a[kTiles * 64];
    __shared__ float s_bias[kBiasTiles * 32];
    __shared__ float s_extra[kExtraFloats];
    __shared__ float s_unscale[kScaleFloats];
    mlp_persistent<Fmt::FP8, true>(args, s_a, s_bias, s_extra, s_unscale);
}

// Caller-given descriptors (mrp_mlp_forward): FP32 rows.
__global__ __launch_bounds__(kMlpBlock) void mlp_bf16_rows_kernel(MlpArgs args) {
    __shared__ bf16x8 s_a[kTiles * 64];
    __shared__ float s_bias[kBiasTiles * 32];
    __shared__ float s_extra[kExtraFloats];
    mlp_persistent<Fmt::BF16, false>(args, s_a, s_bias, s_extra, nullptr);
}

__global__ __launch_bounds__(kMlpBlock) void mlp_fp8_rows_kernel(MlpArgs args) {
    __shared__ long s_a[kTiles * 64];
    __shared__ float s_bias[kBiasTiles * 32];
    __shared__ float s_extra[kExtraFloats];
    __shared__ float s_unscale[kScaleFloats];
    mlp_persistent<Fmt::FP8, false>(args, s_a, s_bias, s_extra, s_unscale);
}

// Per-segment means / maxima from FP32 descriptors (used when the caller has
// no gather-side pool, e.g. mrp_mlp_forward on arbitrary descriptors).
__global__ void pool_kernel(const float* __restrict__ feat, uint32_t n_items, float* __restrict__ pool) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_items * 72) return;
    const size_t item = t / 72;
    const int i = (int)(t % 72);
    const int from[12] = {0, 8, 16, 32, 48, 64, 96, 128, 160, 168, 176, 184};
    const int size[12] = {8, 8, 16, 16, 16, 32, 32, 32, 8, 8, 8, 8};
    const int q = i % 36, blk = q / 3, k = q % 3;
    const float* seg = feat + item * kFeatures + k * kTaps + from[blk];
    float r;
    if (i < 36) {
        float sum = 0.0f;
        for (int j = 0; j < size[blk]; ++j) sum += seg[j];
        r = sum / (float)size[blk];
    } else {
        r = -10000.0f;
        for (int j = 0; j < size[blk]; ++j) r = fmax_(r, seg[j]);
    }
    pool[t] = r;
}

int launch_pool(const float* feat, uint32_t n_items, float* pool, hipStream_t stream) {
    if (n_items == 0) return MRP_OK;
    pool_kernel<<<(unsigned)(((size_t)n_items * 72 + 255) / 256), 256, 0, stream>>>(feat, n_items, pool);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

static int g_num_cus[64] = {};

int launch_mlp(const void* packed, int fmt, const float* feat, const void* tiled, const float* pool,
               const float* gamma, float g, vec3 srp, uint32_t n_items, const uint32_t* n_dev, float* pred,
               hipStream_t stream) {
    if (fmt != (int)Fmt::BF16 && fmt != (int)Fmt::FP8) return fail(MRP_E_INVALID, "unknown packed network format");
    if (n_items == 0) return MRP_OK;
    if (tiled && !pool) return fail(MRP_E_INVALID, "tiled descriptors need the gather's pooled statistics");
    if (n_dev && !pool) return fail(MRP_E_INVALID, "a device-side count needs the gather's pooled statistics");
    if (!tiled && !feat) return fail(MRP_E_INVALID, "no descriptors");
    int dev = 0;
    MRP_HIP(hipGetDevice(&dev));
    int& cus = g_num_cus[dev & 63];
    if (cus == 0) MRP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    float* tmp = nullptr;
    if (!pool) {
        MRP_HIP(hipMalloc(&tmp, (size_t)n_items * 72 * sizeof(float)));
        pool_kernel<<<(unsigned)(((size_t)n_items * 72 + 255) / 256), 256, 0, stream>>>(feat, n_items, tmp);
        pool = tmp;
    }
    const uint32_t tiles = (n_items + 31) / 32;
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)cus, (tiles + kMlpWaves - 1) / kMlpWaves));
    const MlpArgs args{(const uint4*)packed, feat, (const uint4*)tiled, pool, gamma, g, srp, n_items, n_dev, pred};
    if (fmt == (int)Fmt::FP8) {
        if (tiled)
            mlp_fp8_kernel<<<blocks, kMlpBlock, 0, stream>>>(args);
        else
            mlp_fp8_rows_kernel<<<blocks, kMlpBlock, 0, stream>>>(args);
    } else {
        if (tiled)
            mlp_bf16_kernel<<<blocks, kMlpBlock, 0, stream>>>(args);
        else
            mlp_bf16_rows_kernel<<<blocks, kMlpBlock, 0, stream>>>(args);
    }
    hipError_t e = hipGetLastError();
    if (tmp) {
        hipStreamSynchronize(stream);
        hipFree(tmp);
    }
    if (e != hipSuccess) return hip_fail(e, fmt == (int)Fmt::FP8 ? "mlp_fp8_kernel" : "mlp_bf16_kernel");
    return MRP_OK;
}
}  // namespace mrp
