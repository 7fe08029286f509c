// mrp_mlp.h - the MRPNN network (`RadiancePredict` lines 370-486,
// core/radiancePredict.cu) as data: the sequence of squeeze-excitation (SE)
// blocks of the two descriptor streams, then the per-channel head.
//
// Stream state per channel k in {density, transmittance, HG}: x[k] (the gated
// block input, X_ValA/B/C) and y[k] (block outputs, X_ValA2/B2/C2).
//   SE block   : pool mean/max of the three input segments -> 8x8 ReLU ->
//                3 sigmoid gates w_k; x[k][to..] = w_k * in[k][from..]
//                (+ y[k][last..] for a residual block); y[k][to..] = ReLU(W_k x + b_k)
//   ADD        : y[k][to..] += in[k][from..]
//   SHIFT      : y[k][from..] = y[k][to..]; y[k][to..] = in[k][to..]
// Weights are row-major [out][in]; accumulation order follows the reference
// operators (bias first for LinearReLU<S>/_SFL, bias last for the runtime
// LinearReLU), with fused multiply-add as the reference's fast-math build.
#pragma once

#include "mrp_math.h"
#include "mrp_mrpnn.h"
#include "mrp_weights_gen.h"

namespace mrp {

enum : int { OP_SE = 0, OP_SERES = 1, OP_ADD = 2, OP_SHIFT = 3 };

struct StreamOp {
    int op, last, from, to, size, pid;
    int sew0, sew1;   // SE weights (8x8, 3x8)
    int w[3], b[3];   // per-channel S x S weights / biases
};

#define MRP_S3(n) {MRP_W_L##n##W, MRP_W_L_Tr##n##W, MRP_W_L_Hg##n##W}, {MRP_W_L##n##B, MRP_W_L_Tr##n##B, MRP_W_L_Hg##n##B}
#define MRP_D3(n) \
    {MRP_W_LD##n##W, MRP_W_LD_Tr##n##W, MRP_W_LD_Hg##n##W}, {MRP_W_LD##n##B, MRP_W_LD_Tr##n##B, MRP_W_LD_Hg##n##B}
#define MRP_NOW {0, 0, 0}, {0, 0, 0}

// Main stream (taps 0..159) then light stream (taps 160..191), in order.
#define MRP_STREAM_OPS                                                                 \
    {OP_SE, 0, 0, 0, 8, 0, MRP_W_LSE01W, MRP_W_LSE02W, MRP_S3(01)},                    \
        {OP_ADD, 0, 0, 0, 8, 0, 0, 0, MRP_NOW},                                        \
        {OP_SERES, 0, 8, 8, 8, 3, MRP_W_LSE11W, MRP_W_LSE12W, MRP_S3(11)},             \
        {OP_SHIFT, 0, 0, 8, 8, 0, 0, 0, MRP_NOW},                                      \
        {OP_SERES, 0, 16, 16, 16, 6, MRP_W_LSE21W, MRP_W_LSE22W, MRP_S3(21)},          \
        {OP_ADD, 0, 16, 16, 16, 0, 0, 0, MRP_NOW},                                     \
        {OP_SERES, 16, 32, 32, 16, 9, MRP_W_LSE31W, MRP_W_LSE32W, MRP_S3(31)},         \
        {OP_ADD, 0, 32, 32, 16, 0, 0, 0, MRP_NOW},                                     \
        {OP_SERES, 32, 48, 48, 16, 12, MRP_W_LSE41W, MRP_W_LSE42W, MRP_S3(41)},        \
        {OP_SHIFT, 0, 32, 48, 16, 0, 0, 0, MRP_NOW},                                   \
        {OP_SERES, 32, 64, 64, 32, 15, MRP_W_LSE51W, MRP_W_LSE52W, MRP_S3(51)},        \
        {OP_ADD, 0, 64, 64, 32, 0, 0, 0, MRP_NOW},                                     \
        {OP_SERES, 64, 96, 96, 32, 18, MRP_W_LSE61W, MRP_W_LSE62W, MRP_S3(61)},        \
        {OP_ADD, 0, 96, 96, 32, 0, 0, 0, MRP_NOW},                                     \
        {OP_SERES, 96, 128, 128, 32, 21, MRP_W_LSE71W, MRP_W_LSE72W, MRP_S3(71)},      \
        {OP_ADD, 0, 128, 128, 32, 0, 0, 0, MRP_NOW},                                   \
        {OP_SE, 0, 160, 0, 8, 24, MRP_W_LDSE01W, MRP_W_LDSE02W, MRP_D3(01)},           \
        {OP_ADD, 0, 160, 0, 8, 0, 0, 0, MRP_NOW},                                      \
        {OP_SERES, 0, 168, 8, 8, 27, MRP_W_LDSE11W, MRP_W_LDSE12W, MRP_D3(11)},        \
        {OP_ADD, 0, 168, 8, 8, 0, 0, 0, MRP_NOW},                                      \
        {OP_SERES, 8, 176, 16, 8, 30, MRP_W_LDSE21W, MRP_W_LDSE22W, MRP_D3(21)},       \
        {OP_ADD, 0, 176, 16, 8, 0, 0, 0, MRP_NOW},                                     \
        {OP_SERES, 16, 184, 24, 8, 33, MRP_W_LDSE31W, MRP_W_LDSE32W, MRP_D3(31)},      \
    {                                                                                  \
        OP_ADD, 0, 184, 24, 8, 0, 0, 0, MRP_NOW                                        \
    }
constexpr int kStreamOps = 24;
constexpr int kMainOps = 16;  // ops 0..15 are the main stream

// relu(bias + sum in*W) over rows [0,to) of a row-major to x from matrix
MRP_HD void mlp_lin_bias_first(const float* in, float* out, const float* W, const float* B, int from, int to) {
    for (int i = 0; i < to; ++i) {
        float acc = B[i];
        for (int j = 0; j < from; ++j) acc = fmaf(in[j], W[i * from + j], acc);
        out[i] = acc >= 0.0f ? acc : 0.0f;
    }
}

// The complete forward pass for one descriptor (feat = [Xd | Xt | Xh]).
// out[c] = max(exp(P_c) - 1, 0) * srp_c.
MRP_HD void mlp_forward_fp32(const float* w, const float* feat, float gamma, float g, vec3 srp, float* out) {
    const StreamOp ops[kStreamOps] = {MRP_STREAM_OPS};
    float x[3][160], y[3][160], pool[75], comb[128];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 160; ++i) x[k][i] = y[k][i] = 0.0f;
    for (int i = 0; i < 75; ++i) pool[i] = 0.0f;
    pool[72] = g;
    pool[73] = gamma;
    for (int o = 0; o < kStreamOps; ++o) {
        const StreamOp& op = ops[o];
        if (op.op == OP_ADD) {
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < op.size; ++i) y[k][op.to + i] = y[k][op.to + i] + feat[k * kTaps + op.from + i];
            continue;
        }
        if (op.op == OP_SHIFT) {
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < op.size; ++i) {
                    y[k][op.from + i] = y[k][op.to + i];
                    y[k][op.to + i] = feat[k * kTaps + op.to + i];
                }
            continue;
        }
        float t[8], h[8], gate[3];
        for (int k = 0; k < 3; ++k) {
            const float* in = feat + k * kTaps + op.from;
            float sum = 0.0f, mx = -10000.0f;
            for (int i = 0; i < op.size; ++i) {
                sum += in[i];
                mx = fmax_(mx, in[i]);
            }
            pool[op.pid + k] = sum / (float)op.size;
            pool[36 + op.pid + k] = mx;
            t[k] = pool[op.pid + k];
            t[3 + k] = mx;
        }
        t[6] = g;
        t[7] = gamma;
        for (int i = 0; i < 8; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 8; ++j) acc = fmaf(t[j], w[op.sew0 + i * 8 + j], acc);
            h[i] = acc >= 0.0f ? acc : 0.0f;
        }
        for (int i = 0; i < 3; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 8; ++j) acc = fmaf(h[j], w[op.sew1 + i * 8 + j], acc);
            gate[i] = 1.0f / (1.0f + expf_(-acc));
        }
        for (int k = 0; k < 3; ++k) {
            const float* in = feat + k * kTaps + op.from;
            for (int i = 0; i < op.size; ++i) {
                float v = in[i] * gate[k];
                if (op.op == OP_SERES) v = v + y[k][op.last + i];
                x[k][op.to + i] = v;
            }
            mlp_lin_bias_first(&x[k][op.to], &y[k][op.to], w + op.w[k], w + op.b[k], op.size, op.size);
        }
    }
    // last layers of both streams -> comb[0..120)
    const int l8w[3] = {MRP_W_L81W, MRP_W_L_Tr81W, MRP_W_L_Hg81W}, l8b[3] = {MRP_W_L81B, MRP_W_L_Tr81B, MRP_W_L_Hg81B};
    const int d4w[3] = {MRP_W_LD41W, MRP_W_LD_Tr41W, MRP_W_LD_Hg41W},
              d4b[3] = {MRP_W_LD41B, MRP_W_LD_Tr41B, MRP_W_LD_Hg41B};
    (void)kMainOps;
    // the light-stream ops overwrote y[k][0..32); the main stream's y[k][128..160) is intact
    for (int k = 0; k < 3; ++k) mlp_lin_bias_first(&y[k][128], comb + 32 * k, w + l8w[k], w + l8b[k], 32, 32);
    for (int k = 0; k < 3; ++k) mlp_lin_bias_first(&y[k][24], comb + 96 + 8 * k, w + d4w[k], w + d4b[k], 8, 8);
    // head per distinct channel value of srp
    const float sr[3] = {srp.x, srp.y, srp.z};
    float P[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 3; ++c) {
        if (c > 0 && sr[c] == sr[0]) { P[c] = P[0]; continue; }
        if (c == 2 && sr[2] == sr[1]) { P[2] = P[1]; continue; }
        float a[128], b[128], r[16], se1[16], se2[6];
        pool[74] = sr[c];
        for (int i = 0; i < 120; ++i) a[i] = comb[i];
        for (int i = 0; i < 8; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 3; ++j) acc = fmaf(pool[72 + j], w[MRP_W_LGGSW + i * 3 + j], acc);
            acc += w[MRP_W_LGGSB + i];
            a[120 + i] = acc >= 0.0f ? acc : 0.0f;
        }
        for (int i = 0; i < 16; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 75; ++j) acc = fmaf(pool[j], w[MRP_W_LSEFin1W + i * 75 + j], acc);
            se1[i] = acc >= 0.0f ? acc : 0.0f;
        }
        for (int i = 0; i < 6; ++i) {
            float acc = 0.0f;
            for (int j = 0; j < 16; ++j) acc = fmaf(se1[j], w[MRP_W_LSEFin2W + i * 16 + j], acc);
            se2[i] = 1.0f / (1.0f + expf_(-acc));
        }
        for (int i = 0; i < 120; ++i) {
            const int seg = i < 96 ? i / 32 : 3 + (i - 96) / 8;
            a[i] *= se2[seg];
        }
        mlp_lin_bias_first(a, b, w + MRP_W_LC0W, w + MRP_W_LC0B, 128, 128);
        mlp_lin_bias_first(b, a, w + MRP_W_LC1W, w + MRP_W_LC1B, 128, 64);
        mlp_lin_bias_first(a, b, w + MRP_W_LC2W, w + MRP_W_LC2B, 64, 32);
        mlp_lin_bias_first(b, a, w + MRP_W_LXW, w + MRP_W_LXB, 32, 16);
        const int rw0 = MRP_W_LX0W, stride = MRP_W_LX1W - MRP_W_LX0W;  // LXkW, LXkB interleaved
        for (int blk = 0; blk < 6; ++blk) {
            const int o0 = rw0 + (2 * blk) * stride, o1 = rw0 + (2 * blk + 1) * stride;
            mlp_lin_bias_first(a, b, w + o0, w + o0 + 256, 16, 16);
            mlp_lin_bias_first(b, r, w + o1, w + o1 + 256, 16, 16);
            for (int i = 0; i < 16; ++i) a[i] = a[i] + r[i];
        }
        float acc = 0.0f;
        for (int j = 0; j < 16; ++j) acc = fmaf(a[j], w[MRP_W_LX12W + j], acc);
        acc += w[MRP_W_LX12B];
        P[c] = acc >= 0.0f ? acc : 0.0f;
    }
    for (int c = 0; c < 3; ++c) out[c] = fmax_(expf_(P[c]) - 1.0f, 0.0f) * sr[c];
}

// Wave-tiled BF16 descriptors (the renderer's MFMA path: written by the
// gather, read by the MFMA networks).  Descriptors are grouped in tiles of 32;
// for stream k (0..2) and 16-feature chunk c (0..11) of a tile, lane
// l = 32*kg + n owns one 16-byte chunk holding the BF16 features
// 16c + 8h + 4kg + e (h, e = 0..1, 0..3; element 4h + e) of descriptor n: the
// register order of the network's B operands, so one wave load of 1 KB
// fetches a whole 16-feature block of 32 descriptors.  1152 bytes per descriptor.
constexpr int kStreamChunks = kTaps / 16;
constexpr size_t kTileBytes = (size_t)3 * kStreamChunks * 64 * 16;
MRP_HD size_t tile_chunk(uint32_t item, int k, int c, int kg) {
    return (((size_t)(item >> 5) * 3 + (size_t)k) * kStreamChunks + (size_t)c) * 64 + (size_t)(kg * 32) + (item & 31u);
}

#if defined(__HIPCC__)
// MFMA paths (mrp_mlp_mfma.hip); fmt 0: BF16 operands, 1: FP8-E4M3 operands.
// Descriptors: FP32 rows (feat, `tiled` == nullptr) or wave-tiled BF16 (`tiled`).
// With n_dev the count is read on the device and n_items only bounds the grid.
int pack_mlp(const float* weights_host, int fmt, void** packed_dev);
int launch_mlp(const void* packed, int fmt, const float* feat, const void* tiled, const float* pool,
               const float* gamma, float g, vec3 srp, uint32_t n_items, const uint32_t* n_dev, float* pred,
               hipStream_t stream);
// Per-segment means and maxima (72 per descriptor) of n FP32 descriptor rows.
int launch_pool(const float* feat, uint32_t n_items, float* pool, hipStream_t stream);

// BF16 matrix-core training step (mrp_train_mfma.hip): forward and backward of
// a batch into the trainer's record (layout of mrp_train.h), loss and p per
// sample, the batch's loss sum in FP64.
struct TrainMfma;
int train_mfma_create(int max_batch, TrainMfma** out);
void train_mfma_destroy(TrainMfma* m);
int train_mfma_fwd_bwd(TrainMfma* m, const float* w_dev, const float* feat, const float* aux, uint32_t batch,
                       float inv_batch, float* rec, float* loss, float* pred, double* loss_sum,
                       hipStream_t stream);
#endif

}  // namespace mrp
