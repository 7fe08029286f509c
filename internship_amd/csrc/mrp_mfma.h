// mrp_mfma.h - shared pieces of the MRPNN network on gfx950 matrix cores:
// the dense-layer table, the packed-tile layout and the device helpers that
// evaluate one layer for 32 items per wave (D-order registers, see
// mrp_mlp_mfma.hip).  Used by the inference kernels and the training kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>

#include "mrp_math.h"
#include "mrp_weights_gen.h"

namespace mrp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Operand formats of the packed network.
enum class Fmt : int { BF16 = 0, FP8 = 1 };
template <Fmt F> struct OpOf;
template <> struct OpOf<Fmt::BF16> { typedef bf16x8 T; };  // 8 BF16 per lane (16 B)
template <> struct OpOf<Fmt::FP8> { typedef long T; };     // 8 E4M3 per lane (8 B)

// ---------------------------------------------------------------- layer table
enum : int {
    LY_SE0 = 0,         // 12 SE first layers (8x8, no bias)
    LY_SE1 = 12,        // 12 SE second layers (3x8, no bias)
    LY_STREAM = 24,     // 12 blocks x 3 streams (SxS + bias)
    LY_L81 = 60,        // 3 (32x32 + bias)
    LY_LD41 = 63,       // 3 (8x8 + bias)
    LY_FIN1 = 66,       // 16x75, no bias
    LY_FIN2 = 67,       // 6x16, no bias
    LY_LC0 = 68,        // 128x128
    LY_LC1 = 69,        // 64x128
    LY_LC2 = 70,        // 32x64
    LY_LX = 71,         // 16x32
    LY_LXR = 72,        // 12 residual 16x16
    LY_COUNT = 84
};

struct LayerDesc {
    int w, b;      // offsets into the FP32 weight blob (b < 0: no bias)
    int M, K;      // rows, inputs
    int tile;      // first A tile (64 lanes x 8 bf16) in the packed buffer
    int btile;     // first bias tile (2 lane groups x 16 floats) or -1
};

// constexpr, so that the kernels take tile and bias offsets as constants
static constexpr void layer_table(LayerDesc* L) {
    const int se0[12] = {MRP_W_LSE01W, MRP_W_LSE11W, MRP_W_LSE21W, MRP_W_LSE31W, MRP_W_LSE41W, MRP_W_LSE51W,
                         MRP_W_LSE61W, MRP_W_LSE71W, MRP_W_LDSE01W, MRP_W_LDSE11W, MRP_W_LDSE21W, MRP_W_LDSE31W};
    const int se1[12] = {MRP_W_LSE02W, MRP_W_LSE12W, MRP_W_LSE22W, MRP_W_LSE32W, MRP_W_LSE42W, MRP_W_LSE52W,
                         MRP_W_LSE62W, MRP_W_LSE72W, MRP_W_LDSE02W, MRP_W_LDSE12W, MRP_W_LDSE22W, MRP_W_LDSE32W};
    const int sw[12][3] = {{MRP_W_L01W, MRP_W_L_Tr01W, MRP_W_L_Hg01W},   {MRP_W_L11W, MRP_W_L_Tr11W, MRP_W_L_Hg11W},
                           {MRP_W_L21W, MRP_W_L_Tr21W, MRP_W_L_Hg21W},   {MRP_W_L31W, MRP_W_L_Tr31W, MRP_W_L_Hg31W},
                           {MRP_W_L41W, MRP_W_L_Tr41W, MRP_W_L_Hg41W},   {MRP_W_L51W, MRP_W_L_Tr51W, MRP_W_L_Hg51W},
                           {MRP_W_L61W, MRP_W_L_Tr61W, MRP_W_L_Hg61W},   {MRP_W_L71W, MRP_W_L_Tr71W, MRP_W_L_Hg71W},
                           {MRP_W_LD01W, MRP_W_LD_Tr01W, MRP_W_LD_Hg01W}, {MRP_W_LD11W, MRP_W_LD_Tr11W, MRP_W_LD_Hg11W},
                           {MRP_W_LD21W, MRP_W_LD_Tr21W, MRP_W_LD_Hg21W}, {MRP_W_LD31W, MRP_W_LD_Tr31W, MRP_W_LD_Hg31W}};
    const int sb[12][3] = {{MRP_W_L01B, MRP_W_L_Tr01B, MRP_W_L_Hg01B},   {MRP_W_L11B, MRP_W_L_Tr11B, MRP_W_L_Hg11B},
                           {MRP_W_L21B, MRP_W_L_Tr21B, MRP_W_L_Hg21B},   {MRP_W_L31B, MRP_W_L_Tr31B, MRP_W_L_Hg31B},
                           {MRP_W_L41B, MRP_W_L_Tr41B, MRP_W_L_Hg41B},   {MRP_W_L51B, MRP_W_L_Tr51B, MRP_W_L_Hg51B},
                           {MRP_W_L61B, MRP_W_L_Tr61B, MRP_W_L_Hg61B},   {MRP_W_L71B, MRP_W_L_Tr71B, MRP_W_L_Hg71B},
                           {MRP_W_LD01B, MRP_W_LD_Tr01B, MRP_W_LD_Hg01B}, {MRP_W_LD11B, MRP_W_LD_Tr11B, MRP_W_LD_Hg11B},
                           {MRP_W_LD21B, MRP_W_LD_Tr21B, MRP_W_LD_Hg21B}, {MRP_W_LD31B, MRP_W_LD_Tr31B, MRP_W_LD_Hg31B}};
    const int ssize[12] = {8, 8, 16, 16, 16, 32, 32, 32, 8, 8, 8, 8};
    for (int i = 0; i < 12; ++i) {
        L[LY_SE0 + i] = {se0[i], -1, 8, 8, 0, -1};
        L[LY_SE1 + i] = {se1[i], -1, 3, 8, 0, -1};
        for (int k = 0; k < 3; ++k) L[LY_STREAM + 3 * i + k] = {sw[i][k], sb[i][k], ssize[i], ssize[i], 0, -1};
    }
    const int l8w[3] = {MRP_W_L81W, MRP_W_L_Tr81W, MRP_W_L_Hg81W}, l8b[3] = {MRP_W_L81B, MRP_W_L_Tr81B, MRP_W_L_Hg81B};
    const int d4w[3] = {MRP_W_LD41W, MRP_W_LD_Tr41W, MRP_W_LD_Hg41W};
    const int d4b[3] = {MRP_W_LD41B, MRP_W_LD_Tr41B, MRP_W_LD_Hg41B};
    for (int k = 0; k < 3; ++k) {
        L[LY_L81 + k] = {l8w[k], l8b[k], 32, 32, 0, -1};
        L[LY_LD41 + k] = {d4w[k], d4b[k], 8, 8, 0, -1};
    }
    L[LY_FIN1] = {MRP_W_LSEFin1W, -1, 16, 75, 0, -1};
    L[LY_FIN2] = {MRP_W_LSEFin2W, -1, 6, 16, 0, -1};
    L[LY_LC0] = {MRP_W_LC0W, MRP_W_LC0B, 128, 128, 0, -1};
    L[LY_LC1] = {MRP_W_LC1W, MRP_W_LC1B, 64, 128, 0, -1};
    L[LY_LC2] = {MRP_W_LC2W, MRP_W_LC2B, 32, 64, 0, -1};
    L[LY_LX] = {MRP_W_LXW, MRP_W_LXB, 16, 32, 0, -1};
    const int stride = MRP_W_LX1W - MRP_W_LX0W;
    for (int r = 0; r < 12; ++r) L[LY_LXR + r] = {MRP_W_LX0W + r * stride, MRP_W_LX0W + r * stride + 256, 16, 16, 0, -1};
    int tile = 0, btile = 0;
    for (int i = 0; i < LY_COUNT; ++i) {
        const int mt = (L[i].M + 31) / 32, kb = (L[i].K + 15) / 16;
        L[i].tile = tile;
        tile += mt * kb;
        if (L[i].b >= 0) {
            L[i].btile = btile;
            btile += mt;
        }
    }
}

struct LayerTable {
    LayerDesc L[LY_COUNT];
};
static constexpr LayerTable make_layer_table() {
    LayerTable t{};
    layer_table(t.L);
    return t;
}

static constexpr int total_tiles(const LayerDesc* L, int* btiles) {
    int t = 0, b = 0;
    for (int i = 0; i < LY_COUNT; ++i) {
        t += ((L[i].M + 31) / 32) * ((L[i].K + 15) / 16);
        if (L[i].b >= 0) b += (L[i].M + 31) / 32;
    }
    *btiles = b;
    return t;
}

// input feature multiplied by A element e of K block kb for lane group kg
static inline int korder(int kb, int kg, int e) {
    const int j = kb / 2, b = kb % 2;
    return 32 * j + 16 * b + 8 * (e / 4) + 4 * kg + (e % 4);
}
// output row held in register v by lane group kg (MFMA result layout)
MRP_HD int vrow(int v, int kg) { return 8 * (v / 4) + 4 * kg + (v % 4); }

static inline uint16_t to_bf16_rne(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

// Packed buffer: [tiles][64 lanes][8 elements] (BF16: 16 B, FP8: 8 B per lane),
// then [bias tiles][2][16] FP32, then 64 FP32 of small VALU layers: LGGS W (24)
// and b (8), LX12 W (16) and b (1), then 128 FP32 per-layer weight unscales
// (FP8: 2^-sw of layer i at index i; BF16: 1).
constexpr int kExtraFloats = 64, kScaleFloats = 128;
// packed sizes (checked against layer_table) and the persistent kernel's shape
constexpr int kTiles = 150, kBiasTiles = 62, kMlpBlock = 512, kMlpWaves = kMlpBlock / 64;
static_assert(LY_COUNT <= kScaleFloats, "one weight scale per layer");

template <Fmt F>
constexpr size_t tile_bytes() {
    return (size_t)kTiles * 64 * sizeof(typename OpOf<F>::T);
}

// 8 floats -> 8 E4M3 bytes (element e in byte e), hardware rounding.
__device__ __forceinline__ long pack8_fp8(float a0, float a1, float a2, float a3, float a4, float a5, float a6,
                                          float a7) {
    int lo = __builtin_amdgcn_cvt_pk_fp8_f32(a0, a1, 0, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(a2, a3, lo, true);
    int hi = __builtin_amdgcn_cvt_pk_fp8_f32(a4, a5, 0, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(a6, a7, hi, true);
    return (long)(((unsigned long)(unsigned)hi << 32) | (unsigned long)(unsigned)lo);
}
// f(std::integral_constant<int, B + i * D>) for i = 0..N-1: a loop unrolled by
// construction, whatever the compiler's unrolling thresholds (a large body
// left as a loop turns every layer index into a run-time value).
template <int B, int D, class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, B + I * D>{}), ...);
}
template <int B, int N, int D = 1, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl<B, D>(f, std::make_integer_sequence<int, N>{});
}

// n 16-byte words from global memory to LDS by a workgroup of kThreads threads,
// kBatch independent loads in flight per thread (a plain copy loop waits for
// each load before issuing the next).
template <int kThreads, int kBatch = 16>
__device__ __forceinline__ void stage16(const uint4* __restrict__ src, uint4* dst, int n) {
    const int full = n / (kThreads * kBatch) * (kThreads * kBatch);
    for (int i0 = (int)threadIdx.x; i0 < full; i0 += kThreads * kBatch) {
        uint4 t[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) t[j] = src[i0 + j * kThreads];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) dst[i0 + j * kThreads] = t[j];
    }
    for (int i = full + (int)threadIdx.x; i < n; i += kThreads) dst[i] = src[i];
}

// ---------------------------------------------------------------- device helpers
template <Fmt F>
struct Packed {
    const typename OpOf<F>::T* a;  // A tiles
    const float* bias;             // bias tiles
    const float* unscale;          // per-layer weight unscales (FP8)
};

__device__ __forceinline__ bf16x8 to_b(const f32x16& x, int half) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)x[8 * half + e];
    return r;
}

__device__ __forceinline__ long to_fp8(const f32x16& x, int half, float s) {
    const int o = 8 * half;
    return pack8_fp8(x[o] * s, x[o + 1] * s, x[o + 2] * s, x[o + 3] * s, x[o + 4] * s, x[o + 5] * s, x[o + 6] * s,
                     x[o + 7] * s);
}

// max(x, +0) as one integer max on the bits: a negative value or -0 (sign bit
// set: negative as an integer) gives +0.  It differs from x >= 0 ? x : 0 only
// in the sign of a zero, which no later product or sum of the network sees.
__device__ __forceinline__ float relu_bits(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma(long a, long b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int v = 0; v < 16; ++v) z[v] = 0.0f;
    return z;
}

__device__ __forceinline__ void load_bias(const float* bias, int btile, int kg, f32x16& acc) {
    const float4* bp = (const float4*)(bias + ((size_t)btile * 2 + kg) * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 bb = bp[q];
        acc[4 * q] = bb.x;
        acc[4 * q + 1] = bb.y;
        acc[4 * q + 2] = bb.z;
        acc[4 * q + 3] = bb.w;
    }
}

// D (MT tiles) = W x X [+ bias], optional ReLU.  K block kb of the input is
// registers 8*(kb%2)..8*(kb%2)+7 of X[kb/2].
// `layer` indexes the FP8 weight unscales; `d` gives the tiles.
template <Fmt F, int MT, int KB>
__device__ __forceinline__ void dense_d(const Packed<F>& pk, const LayerDesc d, int layer, const f32x16* X,
                                        f32x16 (&D)[MT], bool relu, int lane) {
    const int kg = lane >> 5;
    typename OpOf<F>::T B[KB];
    float unscale = 1.0f;
    if constexpr (F == Fmt::BF16) {
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) B[kb] = to_b(X[kb >> 1], kb & 1);
    } else {
        float amax = 0.0f;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e)  // element e of K block kb is feature 16*kb + 8*(e/4) + ...: skip padding
                if (16 * kb + 8 * (e >> 2) < d.K) amax = __builtin_fmaxf(amax, fabsf(X[kb >> 1][8 * (kb & 1) + e]));
        amax = __builtin_fmaxf(amax, __shfl_xor(amax, 32));  // the item's other lane group
        // amax in [2^(eb-127), 2^(eb-126)): scale 2^(134-eb) puts it in [128, 256)
        uint32_t eb = __float_as_uint(amax) >> 23;
        eb = eb < 8u ? 8u : (eb > 254u ? 254u : eb);
        const float sx = __uint_as_float((261u - eb) << 23);
        unscale = __uint_as_float((eb - 7u) << 23) * pk.unscale[layer];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) B[kb] = to_fp8(X[kb >> 1], kb & 1, sx);
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        f32x16 acc;
        if constexpr (F == Fmt::BF16) {
            if (d.btile >= 0) {
                load_bias(pk.bias, d.btile + mt, kg, acc);
            } else {
                acc = zero16();
            }
        } else {
            acc = zero16();
        }
        const typename OpOf<F>::T* at = pk.a + (size_t)(d.tile + mt * KB) * 64 + lane;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) acc = mfma(at[kb * 64], B[kb], acc);
        // element-wise work only on registers that hold a row < M: register v
        // holds rows mt*32 + 8*(v/4) + 4*kg + v%4, and a padded row has zero
        // weights and bias, so its value is an exact 0 that needs no ReLU
        if constexpr (F == Fmt::FP8) {
            if (d.btile >= 0) {
                f32x16 b;
                load_bias(pk.bias, d.btile + mt, kg, b);
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (mt * 32 + 8 * (v >> 2) < d.M) acc[v] = fmaf(acc[v], unscale, b[v]);
            } else {
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    if (mt * 32 + 8 * (v >> 2) < d.M) acc[v] = acc[v] * unscale;
            }
        }
        if (relu) {
#pragma unroll
            for (int v = 0; v < 16; ++v)
                if (mt * 32 + 8 * (v >> 2) < d.M) acc[v] = relu_bits(acc[v]);
        }
        D[mt] = acc;
    }
}

// 16 features [base, base+16) of a row in D-order into registers v = 8*half..8*half+7
__device__ __forceinline__ void load16(const float* row, int base, int kg, f32x16& x, int half) {
    const float4 lo = *(const float4*)(row + base + 4 * kg);
    const float4 hi = *(const float4*)(row + base + 8 + 4 * kg);
    x[8 * half + 0] = lo.x;
    x[8 * half + 1] = lo.y;
    x[8 * half + 2] = lo.z;
    x[8 * half + 3] = lo.w;
    x[8 * half + 4] = hi.x;
    x[8 * half + 5] = hi.y;
    x[8 * half + 6] = hi.z;
    x[8 * half + 7] = hi.w;
}

__device__ __forceinline__ float sigmoid_(float x) { return 1.0f / (1.0f + expf_(-x)); }

// 4 or 8 BF16 (two per 32-bit word, low half first) -> FP32 registers x[o..]
__device__ __forceinline__ void unpack_bf16(uint32_t w, f32x16& x, int o) {
    x[o] = __uint_as_float(w << 16);
    x[o + 1] = __uint_as_float(w & 0xffff0000u);
}

}  // namespace mrp
