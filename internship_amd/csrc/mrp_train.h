// mrp_train.h - forward and backward of the MRPNN network for one training
// sample, written once for FP32 (the gfx950 training kernel) and FP64 (the CPU
// oracle's gradient).  The reference ships no backward (its training scripts
// train a different network, B11); this is derived from the inference network
// of `RadiancePredict` (core/radiancePredict.cu:370-486), restated in
// mrp_mlp.h, for one colour channel (training samples carry one albedo).
//
// Loss of a sample: relative L2, (p - y)^2 / (y^2 + 0.01), p = max(exp(P)-1, 0) * srp.
//
// The network is a list of 86 dense layers (TrainLayer).  For every layer the
// forward writes its input x and output y into the sample's record and the
// backward writes dL/dz (z = pre-activation).  The weight gradient of a layer
// over a batch is then the product dZ^T X of two record slices, computed by a
// separate GEMM kernel (and by the oracle in FP64).
//
// Q rounds the operands of the products that the matrix-core trainer runs in
// BF16 (layers with lowp set: W and x in the forward, W and dz in dx = W^T dz);
// NoRound leaves them exact.
#pragma once

#include "mrp_mlp.h"

namespace mrp {

enum : int { ACT_NONE = 0, ACT_RELU = 1, ACT_SIGMOID = 2 };

struct TrainLayer {
    int w_off, b_off;   // weight / bias offsets in the 49,673-float blob (b_off < 0: no bias)
    int out, in;
    int act, bias_last; // activation; bias added after (1) or before (0) the sum
    int x_off, y_off, dz_off;  // record slots
    int lowp;           // 1: a matrix-core product in the BF16 trainer (0: LGGS, LX12 on the VALU)
};

struct NoRound {
    template <class T>
    static constexpr T r(T x) { return x; }  // constexpr: host and device
};

// layer ids
constexpr int kSeLayers = 60;            // 12 SE blocks x {SEW0, SEW1, W_A, W_B, W_C}
constexpr int TL_L81 = 60, TL_LD41 = 63;
constexpr int TL_LGGS = 66, TL_FIN1 = 67, TL_FIN2 = 68, TL_LC0 = 69, TL_LC1 = 70, TL_LC2 = 71, TL_LX = 72;
constexpr int TL_LXR = 73;               // LX0..LX11
constexpr int TL_LX12 = 85;
constexpr int kTrainLayers = 86;

// Fill the layer table; returns the record length (floats per sample).
// constexpr, so that kernels can take offsets and sizes as constants.
static constexpr int build_train_layers(TrainLayer* L) {
    const StreamOp ops[kStreamOps] = {MRP_STREAM_OPS};
    int off = 0, n = 0;
    auto add = [&](int w, int b, int out, int in, int act, int bias_last) {
        TrainLayer& l = L[n++];
        l.lowp = w != MRP_W_LGGSW && w != MRP_W_LX12W;
        l.w_off = w;
        l.b_off = b;
        l.out = out;
        l.in = in;
        l.act = act;
        l.bias_last = bias_last;
        l.x_off = off;
        l.y_off = off + in;
        l.dz_off = off + in + out;
        off += in + 2 * out;
    };
    for (int o = 0; o < kStreamOps; ++o) {
        const StreamOp& op = ops[o];
        if (op.op != OP_SE && op.op != OP_SERES) continue;
        add(op.sew0, -1, 8, 8, ACT_RELU, 0);
        add(op.sew1, -1, 3, 8, ACT_SIGMOID, 0);
        for (int k = 0; k < 3; ++k) add(op.w[k], op.b[k], op.size, op.size, ACT_RELU, 0);
    }
    add(MRP_W_L81W, MRP_W_L81B, 32, 32, ACT_RELU, 0);
    add(MRP_W_L_Tr81W, MRP_W_L_Tr81B, 32, 32, ACT_RELU, 0);
    add(MRP_W_L_Hg81W, MRP_W_L_Hg81B, 32, 32, ACT_RELU, 0);
    add(MRP_W_LD41W, MRP_W_LD41B, 8, 8, ACT_RELU, 0);
    add(MRP_W_LD_Tr41W, MRP_W_LD_Tr41B, 8, 8, ACT_RELU, 0);
    add(MRP_W_LD_Hg41W, MRP_W_LD_Hg41B, 8, 8, ACT_RELU, 0);
    add(MRP_W_LGGSW, MRP_W_LGGSB, 8, 3, ACT_RELU, 1);
    add(MRP_W_LSEFin1W, -1, 16, 75, ACT_RELU, 0);
    add(MRP_W_LSEFin2W, -1, 6, 16, ACT_SIGMOID, 0);
    add(MRP_W_LC0W, MRP_W_LC0B, 128, 128, ACT_RELU, 0);
    add(MRP_W_LC1W, MRP_W_LC1B, 64, 128, ACT_RELU, 0);
    add(MRP_W_LC2W, MRP_W_LC2B, 32, 64, ACT_RELU, 0);
    add(MRP_W_LXW, MRP_W_LXB, 16, 32, ACT_RELU, 0);
    const int stride = MRP_W_LX1W - MRP_W_LX0W;  // LXkW, LXkB interleaved
    for (int r = 0; r < 12; ++r) add(MRP_W_LX0W + r * stride, MRP_W_LX0W + r * stride + 256, 16, 16, ACT_RELU, 0);
    add(MRP_W_LX12W, MRP_W_LX12B, 1, 16, ACT_RELU, 1);
    return off;
}

struct TrainTable {
    TrainLayer L[kTrainLayers];
    int rec_floats;
};
static constexpr TrainTable make_train_table() {
    TrainTable t{};
    t.rec_floats = build_train_layers(t.L);
    return t;
}

MRP_HD float t_exp(float x) { return expf_(x); }
MRP_HD float t_fma(float a, float b, float c) { return fmaf(a, b, c); }
#if !defined(__HIP_DEVICE_COMPILE__)
static inline double t_exp(double x) { return exp(x); }
static inline double t_fma(double a, double b, double c) { return a * b + c; }
#endif

// Dense layer forward: records x and y.
template <class T, class Rec, class Q = NoRound>
MRP_HD void t_lin(const float* w, const TrainLayer& l, const T* in, T* out, Rec& rec) {
    for (int j = 0; j < l.in; ++j) rec.put(l.x_off + j, in[j]);
    for (int i = 0; i < l.out; ++i) {
        T acc = (l.b_off >= 0 && !l.bias_last) ? (T)w[l.b_off + i] : (T)0;
        const float* wr = w + l.w_off + i * l.in;
        if (l.lowp) {
            for (int j = 0; j < l.in; ++j) acc = t_fma(Q::r(in[j]), Q::r((T)wr[j]), acc);
        } else {
            for (int j = 0; j < l.in; ++j) acc = t_fma(in[j], (T)wr[j], acc);
        }
        if (l.b_off >= 0 && l.bias_last) acc = acc + (T)w[l.b_off + i];
        if (l.act == ACT_RELU) acc = acc >= (T)0 ? acc : (T)0;
        else if (l.act == ACT_SIGMOID) acc = (T)1 / ((T)1 + t_exp(-acc));
        out[i] = acc;
        rec.put(l.y_off + i, acc);
    }
}

// Dense layer backward: dz = dy * act'(y) (recorded); dx = W^T dz (if dx).
template <class T, class Rec, class Q = NoRound>
MRP_HD void t_lin_back(const float* w, const TrainLayer& l, const T* dy, T* dx, Rec& rec) {
    T dz[128];
    for (int i = 0; i < l.out; ++i) {
        const T y = rec.get(l.y_off + i);
        T d = dy[i];
        if (l.act == ACT_RELU) d = y > (T)0 ? d : (T)0;
        else if (l.act == ACT_SIGMOID) d = d * (y * ((T)1 - y));
        dz[i] = d;
        rec.put(l.dz_off + i, d);
    }
    if (!dx) return;
    for (int j = 0; j < l.in; ++j) {
        T s = (T)0;
        if (l.lowp) {
            for (int i = 0; i < l.out; ++i) s = t_fma(Q::r((T)w[l.w_off + i * l.in + j]), Q::r(dz[i]), s);
        } else {
            for (int i = 0; i < l.out; ++i) s = t_fma((T)w[l.w_off + i * l.in + j], dz[i], s);
        }
        dx[j] = s;
    }
}

// Forward and backward of one sample; returns its loss.  inv_batch scales the
// recorded gradients (the batch loss is the mean).  *pred receives p.
template <class T, class Rec, class Q = NoRound>
MRP_HD T train_sample(const float* w, const TrainLayer* L, const float* feat, float gamma, float g, float srp,
                      float target, T inv_batch, Rec& rec, T* pred) {
    const StreamOp ops[kStreamOps] = {MRP_STREAM_OPS};
    T y[3][160];
    T pool[75];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 160; ++i) y[k][i] = (T)0;
    for (int i = 0; i < 75; ++i) pool[i] = (T)0;
    pool[72] = (T)g;
    pool[73] = (T)gamma;
    pool[74] = (T)srp;
    // ---------------------------------------------------------- streams
    int e = 0;
    for (int o = 0; o < kStreamOps; ++o) {
        const StreamOp& op = ops[o];
        if (op.op == OP_ADD) {
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < op.size; ++i) y[k][op.to + i] = y[k][op.to + i] + (T)feat[k * kTaps + op.from + i];
            continue;
        }
        if (op.op == OP_SHIFT) {
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < op.size; ++i) {
                    y[k][op.from + i] = y[k][op.to + i];
                    y[k][op.to + i] = (T)feat[k * kTaps + op.to + i];
                }
            continue;
        }
        T t[8], h[8], gate[3];
        for (int k = 0; k < 3; ++k) {
            const float* in = feat + k * kTaps + op.from;
            T sum = (T)0, mx = (T)-10000.0f;
            for (int i = 0; i < op.size; ++i) {
                sum = sum + (T)in[i];
                mx = mx > (T)in[i] ? mx : (T)in[i];
            }
            pool[op.pid + k] = sum / (T)op.size;
            pool[36 + op.pid + k] = mx;
            t[k] = pool[op.pid + k];
            t[3 + k] = mx;
        }
        t[6] = (T)g;
        t[7] = (T)gamma;
        t_lin<T, Rec, Q>(w, L[5 * e], t, h, rec);
        t_lin<T, Rec, Q>(w, L[5 * e + 1], h, gate, rec);
        for (int k = 0; k < 3; ++k) {
            T x[32], out[32];
            const float* in = feat + k * kTaps + op.from;
            for (int i = 0; i < op.size; ++i) {
                T v = (T)in[i] * gate[k];
                if (op.op == OP_SERES) v = v + y[k][op.last + i];
                x[i] = v;
            }
            t_lin<T, Rec, Q>(w, L[5 * e + 2 + k], x, out, rec);
            for (int i = 0; i < op.size; ++i) y[k][op.to + i] = out[i];
        }
        ++e;
    }
    // ---------------------------------------------------------- head
    T comb[120], gg[8], se1[16], se2[6], s[128], h0[128], h1[64], h2[32], a[16], b[16], c[16], P;
    for (int k = 0; k < 3; ++k) t_lin<T, Rec, Q>(w, L[TL_L81 + k], &y[k][128], comb + 32 * k, rec);
    for (int k = 0; k < 3; ++k) t_lin<T, Rec, Q>(w, L[TL_LD41 + k], &y[k][24], comb + 96 + 8 * k, rec);
    t_lin<T, Rec, Q>(w, L[TL_LGGS], pool + 72, gg, rec);
    t_lin<T, Rec, Q>(w, L[TL_FIN1], pool, se1, rec);
    t_lin<T, Rec, Q>(w, L[TL_FIN2], se1, se2, rec);
    for (int i = 0; i < 120; ++i) s[i] = comb[i] * se2[i < 96 ? i / 32 : 3 + (i - 96) / 8];
    for (int i = 0; i < 8; ++i) s[120 + i] = gg[i];
    t_lin<T, Rec, Q>(w, L[TL_LC0], s, h0, rec);
    t_lin<T, Rec, Q>(w, L[TL_LC1], h0, h1, rec);
    t_lin<T, Rec, Q>(w, L[TL_LC2], h1, h2, rec);
    t_lin<T, Rec, Q>(w, L[TL_LX], h2, a, rec);
    for (int r = 0; r < 6; ++r) {
        t_lin<T, Rec, Q>(w, L[TL_LXR + 2 * r], a, b, rec);
        t_lin<T, Rec, Q>(w, L[TL_LXR + 2 * r + 1], b, c, rec);
        for (int i = 0; i < 16; ++i) a[i] = a[i] + c[i];
    }
    t_lin<T, Rec, Q>(w, L[TL_LX12], a, &P, rec);
    const T eP = t_exp(P);
    const T p = (eP - (T)1 > (T)0 ? eP - (T)1 : (T)0) * (T)srp;
    *pred = p;
    const T yt = (T)target, den = yt * yt + (T)0.01;
    const T diff = p - yt;
    const T loss = diff * diff / den;
    // ---------------------------------------------------------- backward: head
    T dP = (T)2 * diff / den * inv_batch;
    dP = eP - (T)1 > (T)0 ? dP * (T)srp * eP : (T)0;
    T da[16], dt[16], dc[16], dh2[32], dh1[64], dh0[128], ds[128];
    t_lin_back<T, Rec, Q>(w, L[TL_LX12], &dP, da, rec);
    for (int r = 5; r >= 0; --r) {
        t_lin_back<T, Rec, Q>(w, L[TL_LXR + 2 * r + 1], da, dt, rec);
        t_lin_back<T, Rec, Q>(w, L[TL_LXR + 2 * r], dt, dc, rec);
        for (int i = 0; i < 16; ++i) da[i] = da[i] + dc[i];
    }
    t_lin_back<T, Rec, Q>(w, L[TL_LX], da, dh2, rec);
    t_lin_back<T, Rec, Q>(w, L[TL_LC2], dh2, dh1, rec);
    t_lin_back<T, Rec, Q>(w, L[TL_LC1], dh1, dh0, rec);
    t_lin_back<T, Rec, Q>(w, L[TL_LC0], dh0, ds, rec);
    T dcomb[120], dse2[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0}, dse1[16];
    for (int i = 0; i < 120; ++i) {
        const int seg = i < 96 ? i / 32 : 3 + (i - 96) / 8;
        dcomb[i] = ds[i] * se2[seg];
        dse2[seg] = dse2[seg] + ds[i] * comb[i];
    }
    t_lin_back<T, Rec, Q>(w, L[TL_LGGS], ds + 120, (T*)nullptr, rec);
    t_lin_back<T, Rec, Q>(w, L[TL_FIN2], dse2, dse1, rec);
    t_lin_back<T, Rec, Q>(w, L[TL_FIN1], dse1, (T*)nullptr, rec);
    // ---------------------------------------------------------- backward: streams
    // G[k][i] = dL/dy[k][i] for the current contents of y; the light stream
    // (ops 16..23) and the main stream (ops 0..15) are separate chains.
    T G[3][160];
    for (int chain = 0; chain < 2; ++chain) {
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 160; ++i) G[k][i] = (T)0;
        T dyo[32];
        for (int k = 0; k < 3; ++k) {
            if (chain == 0) {
                t_lin_back<T, Rec, Q>(w, L[TL_LD41 + k], dcomb + 96 + 8 * k, dyo, rec);
                for (int i = 0; i < 8; ++i) G[k][24 + i] = dyo[i];
            } else {
                t_lin_back<T, Rec, Q>(w, L[TL_L81 + k], dcomb + 32 * k, dyo, rec);
                for (int i = 0; i < 32; ++i) G[k][128 + i] = dyo[i];
            }
        }
        const int o_hi = chain == 0 ? kStreamOps - 1 : kMainOps - 1, o_lo = chain == 0 ? kMainOps : 0;
        int se = chain == 0 ? 11 : 7;  // SE-block index of the last SE op of the chain
        for (int o = o_hi; o >= o_lo; --o) {
            const StreamOp& op = ops[o];
            if (op.op == OP_ADD) continue;  // y[to..] += feature: the gradient passes unchanged
            if (op.op == OP_SHIFT) {
                for (int k = 0; k < 3; ++k)
                    for (int i = 0; i < op.size; ++i) {
                        G[k][op.to + i] = G[k][op.from + i];
                        G[k][op.from + i] = (T)0;
                    }
                continue;
            }
            T dgate[3];
            for (int k = 0; k < 3; ++k) {
                T dx[32];
                t_lin_back<T, Rec, Q>(w, L[5 * se + 2 + k], &G[k][op.to], dx, rec);
                const float* in = feat + k * kTaps + op.from;
                T dg = (T)0;
                for (int i = 0; i < op.size; ++i) {
                    G[k][op.to + i] = (T)0;
                    dg = t_fma(dx[i], (T)in[i], dg);
                }
                if (op.op == OP_SERES)
                    for (int i = 0; i < op.size; ++i) G[k][op.last + i] = G[k][op.last + i] + dx[i];
                dgate[k] = dg;
            }
            T dh[8];
            t_lin_back<T, Rec, Q>(w, L[5 * se + 1], dgate, dh, rec);
            t_lin_back<T, Rec, Q>(w, L[5 * se], dh, (T*)nullptr, rec);
            --se;
        }
    }
    return loss;
}

}  // namespace mrp
