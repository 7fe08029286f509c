// mrp_train_mfma.hip - training step of the MRPNN network on gfx950 matrix
// cores: BF16 copies of the FP32 master weights, BF16 activations at every
// product, FP32 accumulation, FP32 records, FP64 weight gradients.
//
// The step is the one of mrp_train.h (derived from `RadiancePredict`,
// core/radiancePredict.cu:370-486, one colour channel per sample) with every
// dense-layer product on v_mfma_f32_32x32x16_bf16, 32 samples per wave in the
// D-order registers of the inference network (mrp_mfma.h):
//   train_pack_*_kernel   - before each step: master weights -> BF16 forward
//                           tiles (the inference layout) and BF16 transposed
//                           tiles (rows = a layer's inputs, K = its outputs),
//                           through element maps built once on the host;
//   train_fwd_mfma_kernel - forward; writes each layer's input x and output y
//                           (FP32, before BF16 rounding) into the trainer's
//                           record rec[slot][sample], the loss and p;
//   train_bwd_mfma_kernel - backward in reverse layer order: dz = dy * act'(y)
//                           with y read back from the record, dz recorded,
//                           dx = W^T dz on the matrix cores.
// The weight gradients dZ^T X then come from the FP64 GEMM of the FP32 path.
// LGGS (8x3) and LX12 (1x16) stay on the VALU in FP32, as in inference.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "mrp_internal.h"
#include "mrp_mfma.h"
#include "mrp_mlp.h"
#include "mrp_train.h"

namespace mrp {

constexpr int kTilesT = 132;          // transposed tiles (layers whose input gradient is needed)
// One tile of 32 samples per wave: a grid-stride loop over tiles would let the
// compiler hoist the record addresses of every layer out of the loop (loop
// invariant), which spills.  Workgroups hold all weight tiles in LDS, so one
// fits a CU; kTrainBlock threads per workgroup.
constexpr int kTrainBlock = 128;
constexpr int kFloatSection = kBiasTiles * 32 + kExtraFloats;

// Transposed tiles for the input gradients: layer ly's W^T (M = its inputs,
// K = its outputs), tile -1 where no input gradient is needed (the first SE
// layers and LSEFin1, whose inputs are pooled statistics).  n_tiles: the
// tiles used.
static constexpr LayerTable make_transposed_table(int* n_tiles) {
    const LayerTable f = make_layer_table();
    LayerTable t{};
    int tt = 0;
    for (int ly = 0; ly < LY_COUNT; ++ly) {
        const bool need_dx = !(ly >= LY_SE0 && ly < LY_SE1) && ly != LY_FIN1;
        t.L[ly] = {f.L[ly].w, -1, f.L[ly].K, f.L[ly].M, need_dx ? tt : -1, -1};
        if (need_dx) tt += ((f.L[ly].K + 31) / 32) * ((f.L[ly].M + 15) / 16);
    }
    *n_tiles = tt;
    return t;
}
static constexpr LayerTable make_transposed_table() {
    int n = 0;
    return make_transposed_table(&n);
}
// forward tiles (the inference table) and transposed tiles, as constants
constexpr LayerTable kTF = make_layer_table();
constexpr LayerTable kTB = make_transposed_table();
static constexpr int transposed_tiles() {
    int n = 0;
    (void)make_transposed_table(&n);
    return n;
}
static_assert(transposed_tiles() == kTilesT, "the transposed tiles must fill the backward kernel's LDS layout exactly");
constexpr TrainTable kTT = make_train_table();  // record slots and sizes, as constants

// record slots of inference layer `ly`: the training layer list orders each SE
// block as {SE first, SE second, stream A, B, C}
MRP_HD int tl_of(int ly) {
    if (ly < LY_SE1) return 5 * (ly - LY_SE0);
    if (ly < LY_STREAM) return 5 * (ly - LY_SE1) + 1;
    if (ly < LY_L81) return 5 * ((ly - LY_STREAM) / 3) + 2 + (ly - LY_STREAM) % 3;
    if (ly < LY_LD41) return TL_L81 + (ly - LY_L81);
    if (ly < LY_FIN1) return TL_LD41 + (ly - LY_LD41);
    if (ly == LY_FIN1) return TL_FIN1;
    if (ly == LY_FIN2) return TL_FIN2;
    if (ly == LY_LC0) return TL_LC0;
    if (ly == LY_LC1) return TL_LC0 + 1;
    if (ly == LY_LC2) return TL_LC0 + 2;
    if (ly == LY_LX) return TL_LX;
    return TL_LXR + (ly - LY_LXR);
}

// ---------------------------------------------------------------- weight copies
__global__ void train_pack_bf16_kernel(const float* __restrict__ w, const int* __restrict__ map, int n,
                                       __bf16* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int src = map[i];
    out[i] = (__bf16)(src >= 0 ? w[src] : 0.0f);
}

__global__ void train_pack_f32_kernel(const float* __restrict__ w, const int* __restrict__ map, int n,
                                      float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int src = map[i];
    out[i] = src >= 0 ? w[src] : 0.0f;
}

// ---------------------------------------------------------------- record access
// A lane's view of the record: sample s (clamped for tail lanes, which write
// nothing).  Slot `slot` of sample s is rec[slot * B + s] (32-bit index: the
// record stays below 2^32 floats).  Row 32j + vrow(v, kg) of a tile is slot
// off + 32j + 8(v/4) + v%4 + 4kg, so the index splits into a part uniform
// over the wave (scalar unit, a constant offset) and the lane's 4 kg B + s.
struct RecLane {
    float* rec;
    uint32_t B, s, base;  // base = 4 kg B + s
    bool valid;
    __device__ __forceinline__ uint32_t at(int off, int j, int v) const {
        return (uint32_t)(off + 32 * j + 8 * (v / 4) + v % 4) * B + base;
    }
    // rows 32j + vrow(v, kg) < n of tile x -> slots off + row
    __device__ __forceinline__ void put(int off, const f32x16& x, int j, int n, int kg) const {
        if (!valid) return;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 32 * j + vrow(v, kg);
            if (r < n) rec[at(off, j, v)] = x[v];
        }
    }
    __device__ __forceinline__ void put1(int slot, float x) const {
        if (valid) rec[(uint32_t)slot * B + s] = x;
    }
    __device__ __forceinline__ float get1(int slot) const { return rec[(uint32_t)slot * B + s]; }
    __device__ __forceinline__ void get(int off, f32x16& y, int j, int n, int kg) const {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 32 * j + vrow(v, kg);
            y[v] = r < n ? rec[at(off, j, v)] : 0.0f;
        }
    }
};

struct TrainMfmaArgs {
    const uint4* packed;    // forward: BF16 tiles, bias tiles, the VALU layers (mrp_mfma.h layout)
    const uint4* packed_t;  // transposed BF16 tiles
    const float* feat;      // B x kFeatures
    const float4* aux;      // gamma, g, srp, target
    uint32_t B;
    float inv_batch;
    float* rec;
    float* loss;
    float* pred;
    double* loss_tiles;     // per 32-sample tile
};

// first feature of each SE block's segment (indexed with compile-time block numbers)
constexpr int kSegFrom[12] = {0, 8, 16, 32, 48, 64, 96, 128, 160, 168, 176, 184};

// the block's input features of stream k, in D-order (rows >= S are zero)
__device__ __forceinline__ void load_seg(const float* frow, int k, int from, int S, int kg, f32x16& xv) {
    xv = zero16();
    const float* srow = frow + k * kTaps + from;
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

// dense layer `ly` forward with its x recorded, and its y for a ReLU layer
// (the two sigmoid layers record their gates, after the sigmoid, themselves)
template <int MT, int KB>
__device__ __forceinline__ void fwd_layer(const Packed<Fmt::BF16>& pk, const RecLane& R, int ly, const f32x16* X,
                                          f32x16 (&Y)[MT], bool relu, int lane) {
    const int kg = (lane >> 5) & 1;  // 0 or 1, visibly to the compiler
    const TrainLayer& l = kTT.L[tl_of(ly)];
#pragma unroll
    for (int j = 0; j < (KB + 1) / 2; ++j) R.put(l.x_off, X[j], j, l.in, kg);
    dense_d<Fmt::BF16, MT, KB>(pk, kTF.L[ly], ly, X, Y, relu, lane);
    if (relu) {
#pragma unroll
        for (int j = 0; j < MT; ++j) R.put(l.y_off, Y[j], j, l.out, kg);
    }
}

// ---------------------------------------------------------------- forward
__global__ __launch_bounds__(kTrainBlock) void train_fwd_mfma_kernel(TrainMfmaArgs a) {
    __shared__ bf16x8 s_a[kTiles * 64];
    __shared__ __align__(16) float s_bias[kBiasTiles * 32];
    __shared__ __align__(16) float s_extra[kExtraFloats];
    {
        constexpr int kA16 = (int)(tile_bytes<Fmt::BF16>() / 16);
        stage16<kTrainBlock>(a.packed, (uint4*)s_a, kA16);
        stage16<kTrainBlock>(a.packed + kA16, (uint4*)s_bias, kBiasTiles * 8);
        stage16<kTrainBlock>(a.packed + kA16 + kBiasTiles * 8, (uint4*)s_extra, kExtraFloats / 4);
    }
    __syncthreads();
    const int lane = (int)__lane_id();
    const int kg = (lane >> 5) & 1;  // 0 or 1, visibly to the compiler
    const Packed<Fmt::BF16> pk{s_a, s_bias, nullptr};
    const int waves = (int)(blockDim.x >> 6);
    const uint32_t n_tiles = (a.B + 31u) / 32u;
    {
        const uint32_t tile = blockIdx.x * waves + (threadIdx.x >> 6);
        if (tile >= n_tiles) return;  // after the staging barrier
        const uint32_t item0 = tile * 32u;
        const uint32_t item = min(item0 + (uint32_t)(lane & 31), a.B - 1u);
        const RecLane R{a.rec, a.B, item, 4u * (uint32_t)kg * a.B + item, item0 + (uint32_t)(lane & 31) < a.B};
        const float4 ax = a.aux[item];
        const float gamma = ax.x, g = ax.y, srp = ax.z, target = ax.w;
        const float* frow = a.feat + (size_t)item * kFeatures;

        f32x16 prev[3] = {zero16(), zero16(), zero16()};
        f32x16 comb[3], lout[3];
        // The head's squeeze-excitation input is pool[3 blk + k] = mean and
        // pool[36 + 3 blk + k] = max of segment blk of stream k, then g, gamma,
        // srp at 72..74.  The pool goes to its record slots (that layer's x) as
        // it is computed, each value by the lane that reads it back at the head.
        const int fin1_x = kTT.L[TL_FIN1].x_off;
        f32x16 xn[3];  // the next block's features, loaded ahead of this block's stores
#pragma unroll
        for (int k = 0; k < 3; ++k) load_seg(frow, k, kSegFrom[0], 8, kg, xn[k]);
        static_for<0, 12>([&](auto blk_c) __attribute__((always_inline)) {
            constexpr int blk = decltype(blk_c)::value;
            constexpr int S = blk < 2 ? 8 : (blk < 5 ? 16 : (blk < 8 ? 32 : 8));
            constexpr bool residual = (blk != 0 && blk != 8);
            constexpr bool rep = (blk == 1 || blk == 4);
            // the block's features, their means and maxima (rows < S: S/2 registers per lane group)
            float mean[3], mx[3];
            f32x16 xs[3] = {xn[0], xn[1], xn[2]};
            if constexpr (blk < 11) {
                constexpr int Sn = blk + 1 < 2 ? 8 : (blk + 1 < 5 ? 16 : (blk + 1 < 8 ? 32 : 8));
#pragma unroll
                for (int k = 0; k < 3; ++k) load_seg(frow, k, kSegFrom[blk + 1], Sn, kg, xn[k]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float sum = 0.0f, m = -10000.0f;
#pragma unroll
                for (int v = 0; v < S / 2; ++v) {
                    sum += xs[k][v];
                    m = fmax_(m, xs[k][v]);
                }
                sum += __shfl_xor(sum, 32);
                mean[k] = sum / (float)S;
                mx[k] = fmax_(m, __shfl_xor(m, 32));
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int r = q == 0 ? 3 * blk + k : 36 + 3 * blk + k, rr = r % 32;
                    if (kg == (rr % 8) / 4) R.put1(fin1_x + r, q == 0 ? mean[k] : mx[k]);
                }
            }
            // squeeze-excitation gates on [means, maxima, g, gamma]
            f32x16 t = zero16();
            if (kg == 0) {
                t[0] = mean[0];
                t[1] = mean[1];
                t[2] = mean[2];
                t[3] = mx[0];
            } else {
                t[0] = mx[1];
                t[1] = mx[2];
                t[2] = g;
                t[3] = gamma;
            }
            f32x16 h[1], o[1];
            fwd_layer<1, 1>(pk, R, LY_SE0 + blk, &t, h, true, lane);
            fwd_layer<1, 1>(pk, R, LY_SE1 + blk, h, o, false, lane);
            f32x16 gs = zero16();
            gs[0] = sigmoid_(o[0][0]);
            gs[1] = sigmoid_(o[0][1]);
            gs[2] = sigmoid_(o[0][2]);
            R.put(kTT.L[tl_of(LY_SE1 + blk)].y_off, gs, 0, 3, kg);  // rows 0..2: lane group 0
            const int src = lane & 31;
            const float gate[3] = {__shfl(gs[0], src), __shfl(gs[1], src), __shfl(gs[2], src)};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const f32x16& xv = xs[k];
                f32x16 x;
#pragma unroll
                for (int v = 0; v < 16; ++v) x[v] = xv[v] * gate[k] + (residual ? prev[k][v] : 0.0f);
                f32x16 y[1];
                if (S == 32) {
                    fwd_layer<1, 2>(pk, R, LY_STREAM + 3 * blk + k, &x, y, true, lane);
                } else {
                    fwd_layer<1, 1>(pk, R, LY_STREAM + 3 * blk + k, &x, y, true, lane);
                }
                if (rep) {
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
                    for (int v = 0; v < 16; ++v) prev[k][v] = y[0][v] + xv[v];
                }
            }
            if (blk == 7) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    f32x16 oc[1];
                    fwd_layer<1, 2>(pk, R, LY_L81 + k, &prev[k], oc, true, lane);
                    comb[k] = oc[0];
                    prev[k] = zero16();
                }
            }
        });
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f32x16 ol[1];
            fwd_layer<1, 1>(pk, R, LY_LD41 + k, &prev[k], ol, true, lane);
            lout[k] = ol[0];
        }
        // ------------------------------------------------ head (the sample's channel)
        f32x16 tp[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) R.get(fin1_x, tp[j], j, 72, kg);
        if (kg == 0) {  // rows 72, 73, 74
            tp[2][4] = g;
            tp[2][5] = gamma;
            tp[2][6] = srp;
        }
        f32x16 s1[1], s2[1];
        fwd_layer<1, 5>(pk, R, LY_FIN1, tp, s1, true, lane);
        fwd_layer<1, 1>(pk, R, LY_FIN2, s1, s2, false, lane);
        f32x16 sg = zero16();
#pragma unroll
        for (int v = 0; v < 4; ++v) sg[v] = sigmoid_(s2[0][v]);
        R.put(kTT.L[TL_FIN2].y_off, sg, 0, 6, kg);
        float s[6];
        const int src0 = lane & 31, src1 = (lane & 31) + 32;
        s[0] = __shfl(sg[0], src0);
        s[1] = __shfl(sg[1], src0);
        s[2] = __shfl(sg[2], src0);
        s[3] = __shfl(sg[3], src0);
        s[4] = __shfl(sg[0], src1);
        s[5] = __shfl(sg[1], src1);
        f32x16 x0[4];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int v = 0; v < 16; ++v) x0[j][v] = comb[j][v] * s[j];
        }
        // LGGS (8x3, bias after the sum) on [g, gamma, srp]: rows 4kg..4kg+3
        {
            const TrainLayer& l = kTT.L[TL_LGGS];
            if (kg == 0) {
                R.put1(l.x_off, g);
                R.put1(l.x_off + 1, gamma);
                R.put1(l.x_off + 2, srp);
            }
            f32x16 lg = zero16();
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 4 * kg + v;
                float acc = 0.0f;
                acc = fmaf(g, s_extra[r * 3 + 0], acc);
                acc = fmaf(gamma, s_extra[r * 3 + 1], acc);
                acc = fmaf(srp, s_extra[r * 3 + 2], acc);
                acc += s_extra[24 + r];
                lg[v] = acc >= 0.0f ? acc : 0.0f;
                x0[3][12 + v] = lg[v];
                x0[3][v] = lout[0][v] * s[3];
                x0[3][4 + v] = lout[1][v] * s[4];
                x0[3][8 + v] = lout[2][v] * s[5];
            }
            R.put(l.y_off, lg, 0, 8, kg);
        }
        f32x16 h0[4], h1[2], h2[1], ha[1];
        fwd_layer<4, 8>(pk, R, LY_LC0, x0, h0, true, lane);
        fwd_layer<2, 8>(pk, R, LY_LC1, h0, h1, true, lane);
        fwd_layer<1, 4>(pk, R, LY_LC2, h1, h2, true, lane);
        fwd_layer<1, 2>(pk, R, LY_LX, h2, ha, true, lane);
        static_for<0, 6>([&](auto r_c) __attribute__((always_inline)) {
            constexpr int r = decltype(r_c)::value;
            f32x16 hb[1], hc[1];
            fwd_layer<1, 1>(pk, R, LY_LXR + 2 * r, ha, hb, true, lane);
            fwd_layer<1, 1>(pk, R, LY_LXR + 2 * r + 1, hb, hc, true, lane);
#pragma unroll
            for (int v = 0; v < 8; ++v) ha[0][v] = ha[0][v] + hc[0][v];
        });
        // LX12 (1x16, bias after the sum)
        const TrainLayer& l12 = kTT.L[TL_LX12];
        R.put(l12.x_off, ha[0], 0, 16, kg);
        float part = 0.0f;
#pragma unroll
        for (int v = 0; v < 8; ++v) part = fmaf(ha[0][v], s_extra[32 + vrow(v, kg)], part);
        part += __shfl_xor(part, 32);
        float P = part + s_extra[48];
        P = P >= 0.0f ? P : 0.0f;
        const float eP = expf_(P);
        const float p = (eP - 1.0f > 0.0f ? eP - 1.0f : 0.0f) * srp;
        const float den = target * target + 0.01f;
        const float diff = p - target;
        const float loss = diff * diff / den;
        if (kg == 0 && R.valid) {
            R.put1(l12.y_off, P);
            a.loss[item] = loss;
            a.pred[item] = p;
        }
        // the tile's loss in FP64, in a fixed order (butterfly over the 32 samples)
        double ls = (kg == 0 && R.valid) ? (double)loss : 0.0;
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) ls += __shfl_xor(ls, o);
        if (lane == 0) a.loss_tiles[tile] = ls;
    }
}

// Sum of the tiles' FP64 losses in a fixed order: 256 strided partial sums, then a tree.
__global__ __launch_bounds__(256) void train_loss_tiles_kernel(const double* __restrict__ v, uint32_t n,
                                                               double* __restrict__ out) {
    __shared__ double part[256];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

// ---------------------------------------------------------------- backward
// On CDNA one counter (vmcnt) tracks a wave's loads and stores together, so a
// load issued after stores cannot be consumed before those stores complete.
// The backward therefore issues the reads of y ahead of stores: in the head,
// each layer's y is loaded before the stores of the layer that precedes it;
// an SE block issues all its reads (five layers' y and its features) at once.

// y tiles of a layer from the record (rows >= out read as 0)
template <int MO>
__device__ __forceinline__ void yload(const RecLane& R, int tl, f32x16 (&Y)[MO], int kg) {
#pragma unroll
    for (int j = 0; j < MO; ++j) R.get(kTT.L[tl].y_off, Y[j], j, kTT.L[tl].out, kg);
}

// dz = dy * act'(y) over rows < out, recorded.
template <int MO>
__device__ __forceinline__ void act_back(const RecLane& R, const TrainLayer& l, const f32x16* dy, const f32x16* y,
                                         f32x16 (&dz)[MO], int kg) {
#pragma unroll
    for (int j = 0; j < MO; ++j) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 32 * j + vrow(v, kg);
            float d = r < l.out ? dy[j][v] : 0.0f;
            if (l.act == ACT_RELU) d = y[j][v] > 0.0f ? d : 0.0f;
            else if (l.act == ACT_SIGMOID) d = d * (y[j][v] * (1.0f - y[j][v]));
            dz[j][v] = d;
        }
        R.put(l.dz_off, dz[j], j, l.out, kg);
    }
}

// backward of inference layer `ly` with its y preloaded: records dz; dX (MI
// tiles over the inputs) = W^T dz
template <int MO, int MI, int KO>
__device__ __forceinline__ void back_layer(const Packed<Fmt::BF16>& pk, const RecLane& R, int ly, const f32x16* dy,
                                           const f32x16* y, f32x16 (&dx)[MI], int lane) {
    f32x16 dz[MO];
    act_back<MO>(R, kTT.L[tl_of(ly)], dy, y, dz, (lane >> 5) & 1);
    dense_d<Fmt::BF16, MI, KO>(pk, kTB.L[ly], ly, dz, dx, false, lane);
}

// What the backward of SE block blk reads: y of its three stream layers and
// its two gate layers, and the block's input features.
struct BlockIn {
    f32x16 ys[3], xv[3], y1, y0;
};

__device__ __forceinline__ void block_load(const RecLane& R, const float* frow, int blk, BlockIn& b, int kg) {
    const int S = blk < 2 ? 8 : (blk < 5 ? 16 : (blk < 8 ? 32 : 8));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const TrainLayer& l = kTT.L[tl_of(LY_STREAM + 3 * blk + k)];
        R.get(l.y_off, b.ys[k], 0, l.out, kg);
        load_seg(frow, k, kSegFrom[blk], S, kg, b.xv[k]);
    }
    R.get(kTT.L[tl_of(LY_SE1 + blk)].y_off, b.y1, 0, 3, kg);
    R.get(kTT.L[tl_of(LY_SE0 + blk)].y_off, b.y0, 0, 8, kg);
}

// One SE block backward: Gp = d loss / d (the block's output window) in, the
// gradient for the previous block's window out (blocks 0 and 8 start a chain).
// All of the block's reads are issued together, before its first store.
template <int blk>
__device__ __forceinline__ void block_back(const Packed<Fmt::BF16>& pk, const RecLane& R, const float* frow,
                                           f32x16 (&Gp)[3], int lane) {
    const int kg = (lane >> 5) & 1;  // 0 or 1, visibly to the compiler
    BlockIn b;
    block_load(R, frow, blk, b, kg);
    constexpr int S = blk < 2 ? 8 : (blk < 5 ? 16 : (blk < 8 ? 32 : 8));
    float dgate[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f32x16 dx[1];
        if (S == 32) {
            back_layer<1, 1, 2>(pk, R, LY_STREAM + 3 * blk + k, &Gp[k], &b.ys[k], dx, lane);
        } else {
            back_layer<1, 1, 1>(pk, R, LY_STREAM + 3 * blk + k, &Gp[k], &b.ys[k], dx, lane);
        }
        float part = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) part = fmaf(dx[0][v], b.xv[k][v], part);
        dgate[k] = part + __shfl_xor(part, 32);
        Gp[k] = dx[0];
    }
    // gates: rows 0..2 of the second SE layer (lane group 0)
    f32x16 dg = zero16();
    dg[0] = dgate[0];
    dg[1] = dgate[1];
    dg[2] = dgate[2];
    f32x16 dh[1];
    back_layer<1, 1, 1>(pk, R, LY_SE1 + blk, &dg, &b.y1, dh, lane);
    f32x16 dz0[1];
    act_back<1>(R, kTT.L[tl_of(LY_SE0 + blk)], dh, &b.y0, dz0, kg);
}

__global__ __launch_bounds__(kTrainBlock) void train_bwd_mfma_kernel(TrainMfmaArgs a) {
    __shared__ bf16x8 s_t[kTilesT * 64];
    __shared__ __align__(16) float s_extra[kExtraFloats];
    {
        stage16<kTrainBlock>(a.packed_t, (uint4*)s_t, kTilesT * 64);
        constexpr int kA16 = (int)(tile_bytes<Fmt::BF16>() / 16);
        stage16<kTrainBlock>(a.packed + kA16 + kBiasTiles * 8, (uint4*)s_extra, kExtraFloats / 4);
    }
    __syncthreads();
    const int lane = (int)__lane_id();
    const int kg = (lane >> 5) & 1;  // 0 or 1, visibly to the compiler
    const Packed<Fmt::BF16> pk{s_t, nullptr, nullptr};
    const int waves = (int)(blockDim.x >> 6);
    const uint32_t n_tiles = (a.B + 31u) / 32u;
    {
        const uint32_t tile = blockIdx.x * waves + (threadIdx.x >> 6);
        if (tile >= n_tiles) return;  // after the staging barrier
        const uint32_t item0 = tile * 32u;
        const uint32_t item = min(item0 + (uint32_t)(lane & 31), a.B - 1u);
        const RecLane R{a.rec, a.B, item, 4u * (uint32_t)kg * a.B + item, item0 + (uint32_t)(lane & 31) < a.B};
        const float4 ax = a.aux[item];
        const float srp = ax.z, target = ax.w;
        const float* frow = a.feat + (size_t)item * kFeatures;

        // ------------------------------------------------ loss -> LX12
        const TrainLayer& l12 = kTT.L[TL_LX12];
        const float P = R.get1(l12.y_off);
        f32x16 ya[1], yb[1];
        yload<1>(R, TL_LXR + 11, ya, kg);
        const float eP = expf_(P);
        const float p = (eP - 1.0f > 0.0f ? eP - 1.0f : 0.0f) * srp;
        const float den = target * target + 0.01f;
        const float diff = p - target;
        float dP = 2.0f * diff / den * a.inv_batch;
        dP = eP - 1.0f > 0.0f ? dP * srp * eP : 0.0f;
        const float dz12 = P > 0.0f ? dP : 0.0f;
        yload<1>(R, TL_LXR + 10, yb, kg);
        if (kg == 0) R.put1(l12.dz_off, dz12);
        f32x16 da[1];
        da[0] = zero16();
#pragma unroll
        for (int v = 0; v < 8; ++v) da[0][v] = s_extra[32 + vrow(v, kg)] * dz12;
        // ------------------------------------------------ residual tower (ya, yb: the pair's y)
        static_for<5, 6, -1>([&](auto r_c) __attribute__((always_inline)) {
            constexpr int r = decltype(r_c)::value;
            f32x16 dt[1], dc[1];
            back_layer<1, 1, 1>(pk, R, LY_LXR + 2 * r + 1, da, ya, dt, lane);
            yload<1>(R, r > 0 ? TL_LXR + 2 * r - 1 : TL_LX, ya, kg);
            back_layer<1, 1, 1>(pk, R, LY_LXR + 2 * r, dt, yb, dc, lane);
            if (r > 0) yload<1>(R, TL_LXR + 2 * r - 2, yb, kg);
#pragma unroll
            for (int v = 0; v < 8; ++v) da[0][v] = da[0][v] + dc[0][v];
        });
        // ------------------------------------------------ head layers (ya: y of LX)
        f32x16 dh2[1], dh1[2], dh0[4], ds[4];
        f32x16 y2[1], y1[2], y0[4];
        yload<1>(R, TL_LC0 + 2, y2, kg);
        back_layer<1, 1, 1>(pk, R, LY_LX, da, ya, dh2, lane);
        yload<2>(R, TL_LC0 + 1, y1, kg);
        back_layer<1, 2, 2>(pk, R, LY_LC2, dh2, y2, dh1, lane);
        yload<4>(R, TL_LC0, y0, kg);
        // what the head input's backward reads: FIN2's gates, the L81 and LD41
        // outputs, LGGS's and FIN1's outputs
        float se2[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) se2[i] = R.get1(kTT.L[TL_FIN2].y_off + i);
        f32x16 cy[3], ly[3], ylg[1], yf1[1];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R.get(kTT.L[TL_L81 + k].y_off, cy[k], 0, 32, kg);
            R.get(kTT.L[TL_LD41 + k].y_off, ly[k], 0, 8, kg);
        }
        yload<1>(R, TL_LGGS, ylg, kg);
        yload<1>(R, TL_FIN1, yf1, kg);
        back_layer<2, 4, 4>(pk, R, LY_LC1, dh1, y1, dh0, lane);
        back_layer<4, 4, 8>(pk, R, LY_LC0, dh0, y0, ds, lane);
        // ------------------------------------------------ head input = [comb * s | LD41 * s | LGGS]
        f32x16 dcomb[3];
        float dse2[6];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float part = 0.0f;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                dcomb[j][v] = ds[j][v] * se2[j];
                part = fmaf(ds[j][v], cy[j][v], part);
            }
            dse2[j] = part + __shfl_xor(part, 32);
        }
        f32x16 dld[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dld[k] = zero16();
            float part = 0.0f;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                dld[k][v] = ds[3][4 * k + v] * se2[3 + k];
                part = fmaf(ds[3][4 * k + v], ly[k][v], part);
            }
            dse2[3 + k] = part + __shfl_xor(part, 32);
        }
        {
            f32x16 dyl = zero16(), dzl[1];
#pragma unroll
            for (int v = 0; v < 4; ++v) dyl[v] = ds[3][12 + v];
            act_back<1>(R, kTT.L[TL_LGGS], &dyl, ylg, dzl, kg);
        }
        {
            // FIN2 rows 0..5: lane group 0 registers 0..3, lane group 1 registers 0..1
            f32x16 dy2 = zero16(), yfin2 = zero16();
            dy2[0] = kg == 0 ? dse2[0] : dse2[4];
            dy2[1] = kg == 0 ? dse2[1] : dse2[5];
            dy2[2] = kg == 0 ? dse2[2] : 0.0f;
            dy2[3] = kg == 0 ? dse2[3] : 0.0f;
            yfin2[0] = kg == 0 ? se2[0] : se2[4];
            yfin2[1] = kg == 0 ? se2[1] : se2[5];
            yfin2[2] = kg == 0 ? se2[2] : 0.0f;
            yfin2[3] = kg == 0 ? se2[3] : 0.0f;
            f32x16 dse1[1], dz1[1];
            back_layer<1, 1, 1>(pk, R, LY_FIN2, &dy2, &yfin2, dse1, lane);
            act_back<1>(R, kTT.L[TL_FIN1], dse1, yf1, dz1, kg);
        }
        // ------------------------------------------------ light stream (blocks 11..8)
        f32x16 Gp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f32x16 o[1];
            back_layer<1, 1, 1>(pk, R, LY_LD41 + k, &dld[k], &ly[k], o, lane);
            Gp[k] = o[0];
        }
        static_for<11, 4, -1>([&](auto b_c) __attribute__((always_inline)) {
            block_back<decltype(b_c)::value>(pk, R, frow, Gp, lane);
        });
        // ------------------------------------------------ main stream (blocks 7..0)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f32x16 o[1];
            back_layer<1, 1, 2>(pk, R, LY_L81 + k, &dcomb[k], &cy[k], o, lane);
            Gp[k] = o[0];
        }
        static_for<7, 8, -1>([&](auto b_c) __attribute__((always_inline)) {
            block_back<decltype(b_c)::value>(pk, R, frow, Gp, lane);
        });
    }
}

// ---------------------------------------------------------------- host side
struct TrainMfma {
    int* map_tiles = nullptr;   // forward tiles then transposed tiles: weight index or -1
    int* map_floats = nullptr;  // bias tiles and the VALU layers
    void* packed = nullptr;     // forward tiles | bias tiles | VALU layers | unscales (mrp_mfma.h layout)
    void* packed_t = nullptr;   // transposed tiles
    double* loss_tiles = nullptr;  // max_batch / 32 (rounded up)
    int max_batch = 0;
};

static const size_t kPackedBytes = tile_bytes<Fmt::BF16>() + (size_t)kFloatSection * 4 + kScaleFloats * 4;

void train_mfma_destroy(TrainMfma* m) {
    if (!m) return;
    (void)hipFree(m->map_tiles);
    (void)hipFree(m->map_floats);
    (void)hipFree(m->packed);
    (void)hipFree(m->packed_t);
    (void)hipFree(m->loss_tiles);
    delete m;
}

int train_mfma_create(int max_batch, TrainMfma** out) {
    *out = nullptr;
    const LayerDesc* F = kTF.L;
    int nb = 0;
    if (total_tiles(F, &nb) != kTiles || nb != kBiasTiles) return fail(MRP_E_STATE, "packed weight tiles do not match the kernel's LDS layout");
    const TrainLayer* TL = kTT.L;
    if ((uint64_t)kTT.rec_floats * (uint64_t)max_batch >= (1ull << 32))
        return fail(MRP_E_INVALID, "BF16 trainer: max_batch too large for 32-bit record indices (at most 1,041,000)");
    // the two layer lists must describe the same products
    for (int ly = 0; ly < LY_COUNT; ++ly) {
        const TrainLayer& l = TL[tl_of(ly)];
        if (l.w_off != F[ly].w || l.b_off != F[ly].b || l.out != F[ly].M || l.in != F[ly].K)
            return fail(MRP_E_STATE, "training layer list does not match the inference layer table");
    }
    // transposed tiles for every layer whose input gradient is needed
    int tt = 0;
    (void)make_transposed_table(&tt);
    const LayerDesc* T = kTB.L;
    if (tt != kTilesT) return fail(MRP_E_STATE, "transposed tile count does not match the kernel's LDS layout");
    std::vector<int> mt((size_t)(kTiles + kTilesT) * 512, -1), mf(kFloatSection, -1);
    for (int ly = 0; ly < LY_COUNT; ++ly) {
        const LayerDesc& d = F[ly];
        const int mtn = (d.M + 31) / 32, kbn = (d.K + 15) / 16;
        for (int m0 = 0; m0 < mtn; ++m0)
            for (int kb = 0; kb < kbn; ++kb)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int m = m0 * 32 + l % 32, f = korder(kb, l / 32, e);
                        if (m < d.M && f < d.K) mt[((size_t)(d.tile + m0 * kbn + kb)) * 512 + l * 8 + e] = d.w + m * d.K + f;
                    }
        if (d.b >= 0)
            for (int m0 = 0; m0 < mtn; ++m0)
                for (int kg = 0; kg < 2; ++kg)
                    for (int v = 0; v < 16; ++v) {
                        const int r = m0 * 32 + vrow(v, kg);
                        if (r < d.M) mf[((size_t)(d.btile + m0) * 2 + kg) * 16 + v] = d.b + r;
                    }
        const LayerDesc& t = T[ly];
        if (t.tile < 0) continue;
        // A^T element: row m = an input of the layer, K index k = an output: W[k][m]
        const int tmn = (t.M + 31) / 32, tkn = (t.K + 15) / 16;
        for (int m0 = 0; m0 < tmn; ++m0)
            for (int kb = 0; kb < tkn; ++kb)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int m = m0 * 32 + l % 32, k = korder(kb, l / 32, e);
                        if (m < t.M && k < t.K)
                            mt[((size_t)(kTiles + t.tile + m0 * tkn + kb)) * 512 + l * 8 + e] = d.w + k * d.K + m;
                    }
    }
    int* extra = mf.data() + kBiasTiles * 32;
    for (int i = 0; i < 24; ++i) extra[i] = MRP_W_LGGSW + i;
    for (int i = 0; i < 8; ++i) extra[24 + i] = MRP_W_LGGSB + i;
    for (int i = 0; i < 16; ++i) extra[32 + i] = MRP_W_LX12W + i;
    extra[48] = MRP_W_LX12B;

    TrainMfma* m = new (std::nothrow) TrainMfma();
    if (!m) return fail(MRP_E_NOMEM, "host allocation failed");
    m->max_batch = max_batch;
    bool ok = hipMalloc(&m->map_tiles, mt.size() * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&m->map_floats, mf.size() * sizeof(int)) == hipSuccess;
    ok = ok && hipMalloc(&m->packed, kPackedBytes) == hipSuccess;
    ok = ok && hipMalloc(&m->packed_t, (size_t)kTilesT * 64 * 16) == hipSuccess;
    ok = ok && hipMalloc(&m->loss_tiles, (size_t)(max_batch + 31) / 32 * sizeof(double)) == hipSuccess;
    if (!ok) {
        train_mfma_destroy(m);
        return fail(MRP_E_NOMEM, "trainer allocation failed");
    }
    std::vector<float> ones(kScaleFloats, 1.0f);  // per-layer unscales (used by FP8 only)
    ok = hipMemcpy(m->map_tiles, mt.data(), mt.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(m->map_floats, mf.data(), mf.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy((char*)m->packed + tile_bytes<Fmt::BF16>() + (size_t)kFloatSection * 4, ones.data(),
                         kScaleFloats * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        train_mfma_destroy(m);
        return fail(MRP_E_HIP, "trainer initialisation failed");
    }
    *out = m;
    return MRP_OK;
}

int train_mfma_fwd_bwd(TrainMfma* m, const float* w, const float* feat, const float* aux, uint32_t B,
                       float inv_batch, float* rec, float* loss, float* pred, double* loss_sum,
                       hipStream_t stream) {
    if (B == 0 || (int)B > m->max_batch) return fail(MRP_E_INVALID, "batch must be in [1, max_batch]");
    // BF16 copies of the master weights
    __bf16* tiles_f = (__bf16*)m->packed;
    __bf16* tiles_t = (__bf16*)m->packed_t;
    train_pack_bf16_kernel<<<(kTiles * 512 + 255) / 256, 256, 0, stream>>>(w, m->map_tiles, kTiles * 512, tiles_f);
    train_pack_bf16_kernel<<<(kTilesT * 512 + 255) / 256, 256, 0, stream>>>(w, m->map_tiles + kTiles * 512,
                                                                            kTilesT * 512, tiles_t);
    train_pack_f32_kernel<<<(kFloatSection + 255) / 256, 256, 0, stream>>>(
        w, m->map_floats, kFloatSection, (float*)((char*)m->packed + tile_bytes<Fmt::BF16>()));
    MRP_HIP(hipGetLastError());
    const TrainMfmaArgs args{(const uint4*)m->packed, (const uint4*)m->packed_t, feat, (const float4*)aux,
                             B, inv_batch, rec, loss, pred, m->loss_tiles};
    const uint32_t tiles = (B + 31u) / 32u, waves = kTrainBlock / 64;
    const uint32_t blocks = (tiles + waves - 1) / waves;
    train_fwd_mfma_kernel<<<blocks, kTrainBlock, 0, stream>>>(args);
    MRP_HIP(hipGetLastError());
    train_bwd_mfma_kernel<<<blocks, kTrainBlock, 0, stream>>>(args);
    MRP_HIP(hipGetLastError());
    train_loss_tiles_kernel<<<1, 256, 0, stream>>>(m->loss_tiles, tiles, loss_sum);
    MRP_HIP(hipGetLastError());
    return MRP_OK;
}

}  // namespace mrp
