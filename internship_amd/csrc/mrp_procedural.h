// mrp_procedural.h - seeded procedural clouds used as synthetic inputs.
//
// The reference's procedural volume (`GenerateModel`, genall.cu:32-67) calls a
// `sin*cos*sin` stub (core/perlin_noise.cuh:4-7) whose .cu does not compile, so
// the clouds here are this project's own: Perlin's improved gradient noise
// (K. Perlin, "Improving Noise", SIGGRAPH 2002: quintic fade, 12 edge
// gradients, 256-entry permutation), summed as fractal Brownian motion,
// thresholded, remapped to [0,1] and shaped by a spherical falloff.
//
//   n(p)      = sum_{o<octaves} gain^o * noise(p * freq * lacunarity^o) / sum gain^o
//   v         = (n + 1) / 2
//   d         = clamp((v - threshold) / (1 - threshold), 0, 1)
//   falloff   = clamp((0.5 - |p - c|) / 0.2, 0, 1),  c = (0.5,0.5,0.5)
//   density   = d * falloff
//
// p is the voxel centre (i+0.5)/res.  The permutation is a Fisher-Yates
// shuffle of 0..255 driven by Philox(seed).  Only IEEE basic operations are
// used, so the GPU fill and the CPU fill produce the same bits.
#pragma once

#include "mrp_math.h"

namespace mrp {

struct CloudParams {
    uint64_t seed;
    int octaves;
    float base_freq, lacunarity, gain, threshold;
};

// perm must hold 512 ints (the 256-entry permutation repeated).
static inline void cloud_permutation(uint64_t seed, int* perm) {
    for (int i = 0; i < 256; ++i) perm[i] = i;
    Rng r = rng_make(0u, 0u, 7u, seed);
    for (int i = 255; i > 0; --i) {
        const float u = rng_next(&r);  // (0,1]
        int j = (int)(u * (float)(i + 1)) - 1;
        if (j < 0) j = 0;
        if (j > i) j = i;
        const int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    for (int i = 0; i < 256; ++i) perm[256 + i] = perm[i];
}

MRP_HD float perlin_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
MRP_HD float perlin_lerp(float t, float a, float b) { return a + t * (b - a); }
MRP_HD float perlin_grad(int h, float x, float y, float z) {
    h &= 15;
    const float u = h < 8 ? x : y;
    const float v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
MRP_HD float perlin_noise(const int* p, float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int X = ((int)fx) & 255, Y = ((int)fy) & 255, Z = ((int)fz) & 255;
    x -= fx;
    y -= fy;
    z -= fz;
    const float u = perlin_fade(x), v = perlin_fade(y), w = perlin_fade(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    const int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return perlin_lerp(w,
                       perlin_lerp(v, perlin_lerp(u, perlin_grad(p[AA], x, y, z), perlin_grad(p[BA], x - 1.0f, y, z)),
                                   perlin_lerp(u, perlin_grad(p[AB], x, y - 1.0f, z),
                                               perlin_grad(p[BB], x - 1.0f, y - 1.0f, z))),
                       perlin_lerp(v,
                                   perlin_lerp(u, perlin_grad(p[AA + 1], x, y, z - 1.0f),
                                               perlin_grad(p[BA + 1], x - 1.0f, y, z - 1.0f)),
                                   perlin_lerp(u, perlin_grad(p[AB + 1], x, y - 1.0f, z - 1.0f),
                                               perlin_grad(p[BB + 1], x - 1.0f, y - 1.0f, z - 1.0f))));
}

MRP_HD float cloud_density(const int* perm, const CloudParams& c, int res, int ix, int iy, int iz) {
    const float inv = 1.0f / (float)(2 * res);
    const float px = (float)(2 * ix + 1) * inv, py = (float)(2 * iy + 1) * inv, pz = (float)(2 * iz + 1) * inv;
    float sum = 0.0f, norm = 0.0f, amp = 1.0f, freq = c.base_freq;
    for (int o = 0; o < c.octaves; ++o) {
        sum += amp * perlin_noise(perm, px * freq, py * freq, pz * freq);
        norm += amp;
        amp *= c.gain;
        freq *= c.lacunarity;
    }
    const float n = norm > 0.0f ? sum / norm : 0.0f;
    const float v = (n + 1.0f) * 0.5f;
    float d = (v - c.threshold) / (1.0f - c.threshold);
    d = fmin_(fmax_(d, 0.0f), 1.0f);
    const float dx = px - 0.5f, dy = py - 0.5f, dz = pz - 0.5f;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
    const float fall = fmin_(fmax_((0.5f - r) / 0.2f, 0.0f), 1.0f);
    return d * fall;
}

}  // namespace mrp
