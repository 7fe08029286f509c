// mrp_math.h - IEEE-only scalar math shared by the gfx950 kernels and the CPU oracle.
//
// Every function here is built from +, -, *, /, sqrt (all correctly rounded in
// IEEE binary32 on both gfx950 and x86-64), floor, fabs, fmin/fmax and integer
// bit manipulation.  No library transcendental is called, so the same source
// compiled by hipcc (device) and by g++ (host, -ffp-contract=off) returns the
// same bits for the same input.  That is what lets a path traced on the GPU and
// the same path traced by the CPU oracle take identical branch decisions
// (S/SMax > xi), which a 1-ulp difference in log() would otherwise flip.
//
// Methods (all published, coefficients are plain Taylor coefficients):
//   log  : x = 2^e * m, m in [sqrt(1/2), sqrt(2)), f = m-1; log(1+f) = f P(f)
//          with a fitted degree-9 polynomial (division-free); ln2 split in two
//          parts (Cody & Waite 1980).
//   fused multiply-add (fmaf) is correctly rounded on both targets, so explicit
//   fmaf keeps CPU and GPU bit-identical (the oracle is built with -mfma).
//   exp  : n = round(x log2 e); r = x - n ln2 (Cody-Waite two-part constant);
//          e^r by its Taylor series to degree 7; 2^n by exponent construction.
//   sin/cos : quadrant reduction by pi/2 (three-part Cody-Waite constant),
//          Taylor series on [-pi/4, pi/4].
//   atan : reduction |x|>1 -> pi/2 - atan(1/x); |x|>tan(pi/12) ->
//          pi/6 + atan((sqrt3 x - 1)/(x + sqrt3)); odd Taylor series to x^13.
//   acos : acos(x) = 2 atan(sqrt((1-x)/(1+x))).
//   Philox4x32-10 : Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//          easy as 1, 2, 3", SC'11 (multipliers 0xD2511F53/0xCD9E8D57, Weyl
//          constants 0x9E3779B9/0xBB67AE85).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MRP_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define MRP_HD static inline
#endif

namespace mrp {

// ---------------------------------------------------------------- bit casts
MRP_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
MRP_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

MRP_HD float f_inf() { return u2f(0x7f800000u); }
MRP_HD float f_nan() { return u2f(0x7fc00000u); }

MRP_HD float fmin_(float a, float b) { return a < b ? a : b; }
MRP_HD float fmax_(float a, float b) { return a > b ? a : b; }
MRP_HD float fabs_(float a) { return u2f(f2u(a) & 0x7fffffffu); }

// ---------------------------------------------------------------- vec3
struct vec3 {
    float x, y, z;
};
MRP_HD vec3 v3(float x, float y, float z) {
    vec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
MRP_HD vec3 operator+(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
MRP_HD vec3 operator-(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
MRP_HD vec3 operator*(vec3 a, vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
MRP_HD vec3 operator*(vec3 a, float b) { return v3(a.x * b, a.y * b, a.z * b); }
MRP_HD vec3 operator/(vec3 a, float b) { return v3(a.x / b, a.y / b, a.z / b); }
MRP_HD vec3 operator+(vec3 a, float b) { return v3(a.x + b, a.y + b, a.z + b); }
MRP_HD vec3 operator-(vec3 a, float b) { return v3(a.x - b, a.y - b, a.z - b); }
MRP_HD vec3 operator-(vec3 a) { return v3(-a.x, -a.y, -a.z); }
MRP_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MRP_HD vec3 cross(vec3 a, vec3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
MRP_HD vec3 vmin(vec3 a, vec3 b) { return v3(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
MRP_HD vec3 vmax(vec3 a, vec3 b) { return v3(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }
MRP_HD float length(vec3 a) { return sqrtf(dot(a, a)); }
MRP_HD float distance(vec3 a, vec3 b) { return length(a - b); }
MRP_HD vec3 normalize(vec3 v) {
    const float inv = 1.0f / sqrtf(dot(v, v));
    return v3(v.x * inv, v.y * inv, v.z * inv);
}
// a + v*(b-a) with one rounding of the product-sum: the blend used for
// texture interpolation and running means (equal to a*(1-v) + b*v up to
// rounding; exact for a == b).
MRP_HD float lerp(float a, float b, float v) { return fmaf(v, b - a, a); }
MRP_HD vec3 lerp(vec3 a, vec3 b, float v) { return v3(lerp(a.x, b.x, v), lerp(a.y, b.y, v), lerp(a.z, b.z, v)); }

// ---------------------------------------------------------------- log
MRP_HD float logf_(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -f_inf() : f_nan();
    if (x == f_inf()) return x;
    int e = 0;
    uint32_t b = f2u(x);
    if (b < 0x00800000u) {  // subnormal: scale by 2^24
        x = x * 16777216.0f;
        b = f2u(x);
        e = -24;
    }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;  // exact (Sterbenz), f in [sqrt(1/2)-1, sqrt(2)-1]
    // log(1+f) = f * P(f), P a degree-9 polynomial fitted to log1p(f)/f on
    // this interval (weighted least squares on Chebyshev nodes; 1.6 ulp max),
    // evaluated by Horner with fused multiply-add: no division.
    float p = -7.830642164e-02f;
    p = fmaf(p, f, 1.301906109e-01f);
    p = fmaf(p, f, -1.321229190e-01f);
    p = fmaf(p, f, 1.416073740e-01f);
    p = fmaf(p, f, -1.661034077e-01f);
    p = fmaf(p, f, 2.000320554e-01f);
    p = fmaf(p, f, -2.500157058e-01f);
    p = fmaf(p, f, 3.333330154e-01f);
    p = fmaf(p, f, -4.999998808e-01f);
    p = fmaf(p, f, 1.0f);
    const float lm = f * p;
    const float fe = (float)e;
    // ln2 = 0.693145751953125 (exactly representable, 16 significant bits) + 1.42860677e-06
    return fe * 0.693145751953125f + (lm + fe * 1.42860677e-06f);
}

// ---------------------------------------------------------------- exp
MRP_HD float exp2i_(int n) {  // 2^n for n in [-126, 127]
    return u2f((uint32_t)(n + 127) << 23);
}
MRP_HD float expf_(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return f_inf();
    if (x < -103.972084f) return 0.0f;
    const float kf = floorf(x * 1.44269502f + 0.5f);
    const int k = (int)kf;
    // ln2 = 0.693359375 (9 significant bits: k*C1 is exact) - 2.12194440e-4
    const float r = (x - kf * 0.693359375f) + kf * 2.12194440e-4f;
    // e^r, |r| <= 0.35, Taylor to r^7
    float p = 1.98412701e-4f;          // 1/5040
    p = p * r + 1.38888892e-3f;        // 1/720
    p = p * r + 8.33333377e-3f;        // 1/120
    p = p * r + 4.16666679e-2f;        // 1/24
    p = p * r + 1.66666672e-1f;        // 1/6
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) return p * exp2i_(127) * exp2i_(k - 127);
    if (k < -126) return p * exp2i_(k + 126) * exp2i_(-126);
    return p * exp2i_(k);
}

// ---------------------------------------------------------------- sin / cos
MRP_HD void sincos_poly_(float r, float* s, float* c) {
    const float z = r * r;
    // sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9!
    float ps = 2.75573188e-6f;
    ps = ps * z - 1.98412701e-4f;
    ps = ps * z + 8.33333377e-3f;
    ps = ps * z - 1.66666672e-1f;
    *s = r + r * (z * ps);
    // cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
    float pc = -2.75573188e-7f;
    pc = pc * z + 2.48015876e-5f;
    pc = pc * z - 1.38888892e-3f;
    pc = pc * z + 4.16666679e-2f;
    pc = pc * z - 0.5f;
    *c = 1.0f + z * pc;
}
// Valid for |x| < 2^15 (ample for angles built from uniforms in (0,1]).
MRP_HD void sincosf_(float x, float* sn, float* cs) {
    const float kf = floorf(x * 0.636619747f + 0.5f);
    const int k = (int)kf;
    // pi/2 = 1.5703125 + 4.83870506e-4 - 4.37113883e-8 (first two parts short,
    // so k*C1 and k*C2 are exact for |k| < 2^11)
    float r = x - kf * 1.5703125f;
    r = r - kf * 4.83870506e-4f;
    r = r + kf * 4.37113883e-8f;
    float s, c;
    sincos_poly_(r, &s, &c);
    switch (k & 3) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

// ---------------------------------------------------------------- atan / acos
MRP_HD float atanf_(float x) {
    if (x != x) return x;
    const bool neg = x < 0.0f;
    float a = fabs_(x);
    float base = 0.0f;
    bool inv = false;
    if (a > 1.0f) {
        a = 1.0f / a;  // inf -> 0
        inv = true;
    }
    if (a > 0.267949194f) {  // tan(pi/12)
        a = (a * 1.73205078f - 1.0f) / (a + 1.73205078f);
        base = 0.523598790f;  // pi/6
    }
    const float z = a * a;
    float p = 7.69230798e-2f;       // 1/13
    p = p * z - 9.09090936e-2f;     // -1/11
    p = p * z + 0.111111112f;       // 1/9
    p = p * z - 0.142857149f;       // -1/7
    p = p * z + 0.200000003f;       // 1/5
    p = p * z - 0.333333343f;       // -1/3
    float r = base + (a + a * (z * p));
    if (inv) r = 1.57079637f - r;
    return neg ? -r : r;
}
MRP_HD float acosf_(float x) {
    x = fmin_(fmax_(x, -1.0f), 1.0f);
    return 2.0f * atanf_(sqrtf((1.0f - x) / (1.0f + x)));
}
// atan2 by quadrant from atanf_: (-pi, pi]; atan2(0, 0) = 0, atan2(+-0, x<0) = +pi.
MRP_HD float atan2f_(float y, float x) {
    if (x > 0.0f) return atanf_(y / x);
    if (x < 0.0f) return y >= 0.0f ? atanf_(y / x) + 3.14159274f : atanf_(y / x) - 3.14159274f;
    return y > 0.0f ? 1.57079637f : (y < 0.0f ? -1.57079637f : 0.0f);
}

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
    uint32_t v[4];
};
MRP_HD uint32_t pick4_(const u32x4& b, uint32_t w) {
    return w == 0u ? b.v[0] : (w == 1u ? b.v[1] : (w == 2u ? b.v[2] : b.v[3]));
}
MRP_HD void mulhilo_(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}
MRP_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo_(0xD2511F53u, c.v[0], &hi0, &lo0);
        mulhilo_(0xCD9E8D57u, c.v[2], &hi1, &lo1);
        u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
// 32 random bits -> float in (0, 1]: ((x >> 8) + 1) * 2^-24, the (0,1] range of
// the draws the reference path tracer consumes.
MRP_HD float u01_(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f; }

// One counter-based stream per (pixel, sample): counter = (stream_id, sample,
// block, tag), key = (seed lo, seed hi).  Draw n comes from block n/4, word n%4.
struct Rng {
    uint32_t id, sample, tag, k0, k1;
    uint32_t n;
    u32x4 buf;
};
MRP_HD Rng rng_make(uint32_t id, uint32_t sample, uint32_t tag, uint64_t seed) {
    Rng r;
    r.id = id;
    r.sample = sample;
    r.tag = tag;
    r.k0 = (uint32_t)seed;
    r.k1 = (uint32_t)(seed >> 32);
    r.n = 0;
    r.buf.v[0] = r.buf.v[1] = r.buf.v[2] = r.buf.v[3] = 0u;
    return r;
}
// Resume a stream at draw n (the state rng_next leaves after n draws).
MRP_HD void rng_seek(Rng* r, uint32_t n) {
    r->n = n;
    if (n & 3u) {
        u32x4 c;
        c.v[0] = r->id;
        c.v[1] = r->sample;
        c.v[2] = n >> 2;
        c.v[3] = r->tag;
        r->buf = philox4x32_10(c, r->k0, r->k1);
    }
}
MRP_HD float rng_next(Rng* r) {
    const uint32_t w = r->n & 3u;
    if (w == 0u) {
        u32x4 c;
        c.v[0] = r->id;
        c.v[1] = r->sample;
        c.v[2] = r->n >> 2;
        c.v[3] = r->tag;
        r->buf = philox4x32_10(c, r->k0, r->k1);
    }
    r->n += 1u;
    return u01_(pick4_(r->buf, w));
}

// The same stream read through a ring of kRingWords generated words, so that a
// walk can make all the words of a window of draws at one site (rng_fill) and
// read them by index (rng_word): a wave then runs the cipher once per missing
// block, not once per draw site that some lane reaches at a block boundary,
// and a rewind inside the window (the batched trackers' speculation) costs
// nothing.  Word i sits at w[(i % kRingWords) * kStride]: kStride 64 places
// the 64 lanes of a wave side by side in LDS (a wave's reads of any words hit
// 64 different banks), 1 is a private array.  hi, a multiple of 4, ends the
// generated words; the ring holds words [hi - kRingWords, hi).  A seek may go
// forward anywhere, or back over words taken since the ring last restarted.
constexpr uint32_t kRingWords = 16u;
template <uint32_t kStride>
struct RingRng {
    uint32_t* w;
    uint32_t id, sample, tag, k0, k1;
    uint32_t n, hi;
};
template <uint32_t kStride>
MRP_HD RingRng<kStride> ring_make(uint32_t* w, uint32_t id, uint32_t sample, uint32_t tag, uint64_t seed) {
    RingRng<kStride> r;
    r.w = w;
    r.id = id;
    r.sample = sample;
    r.tag = tag;
    r.k0 = (uint32_t)seed;
    r.k1 = (uint32_t)(seed >> 32);
    r.n = 0u;
    r.hi = 0u;
    return r;
}
// Makes the words below `end`; with end - n <= kRingWords - 3 every word not
// yet taken stays in the ring.
template <uint32_t kStride>
MRP_HD void rng_fill(RingRng<kStride>* r, uint32_t end) {
    while (r->hi < end) {
        u32x4 c;
        c.v[0] = r->id;
        c.v[1] = r->sample;
        c.v[2] = r->hi >> 2;
        c.v[3] = r->tag;
        const u32x4 o = philox4x32_10(c, r->k0, r->k1);
        uint32_t* d = r->w + (r->hi & (kRingWords - 1u)) * kStride;
        d[0] = o.v[0];
        d[kStride] = o.v[1];
        d[2 * kStride] = o.v[2];
        d[3 * kStride] = o.v[3];
        r->hi += 4u;
    }
}
// Draw i (made by rng_fill, still in the ring), without taking it.
template <uint32_t kStride>
MRP_HD float rng_word(const RingRng<kStride>& r, uint32_t i) {
    return u01_(r.w[(i & (kRingWords - 1u)) * kStride]);
}
template <uint32_t kStride>
MRP_HD float rng_next(RingRng<kStride>* r) {
    rng_fill(r, r->n + 1u);
    const float u = rng_word(*r, r->n);
    r->n += 1u;
    return u;
}
template <uint32_t kStride>
MRP_HD void rng_seek(RingRng<kStride>* r, uint32_t n) {
    if (n > r->hi || n + kRingWords < r->hi) r->hi = n & ~3u;  // outside the ring: restart at n's block
    r->n = n;
}
#if defined(__HIPCC__)
// This thread's ring in a workgroup's LDS array of blockDim.x * kRingWords
// words: word i of lane l of wave v at [(v * kRingWords + i % kRingWords) * 64 + l].
__device__ __forceinline__ uint32_t* lds_ring(uint32_t* ring) {
    return ring + (threadIdx.x >> 6) * (kRingWords * 64u) + (threadIdx.x & 63u);
}
#endif

}  // namespace mrp
