// mrp_hglut.h - cone-averaged Henyey-Greenstein table (`Fill_Hg`,
// core/volume.cu:1082-1160), shared by the gfx950 kernel and the CPU oracle.
//
// Texel (i, j) of the 512 x 512 table holds the average of the HG phase over a
// cone of directions: i indexes cos(theta) in [-1, 1], j the cone half-angle in
// [0, 60 deg].  1024 samples per texel mix two estimators with a
// multiple-importance weight: HG-lobe samples scored by the chord length of the
// cone's ball, and uniform samples inside the ball scored by HG.  The sample
// points are scrambled Halton points (PBRT's ScrambledRadicalInverse, bases 2,
// 3, 5, 11) shifted by one random offset per texel (Cranley-Patterson rotation).
#pragma once

#include "mrp_medium.h"
#include "mrp_tables_gen.h"

namespace mrp {

constexpr int kHgLutSize = 512;
constexpr int kHgLutSamples = 1024;

// Scrambled radical inverse of `a` in prime base `base`, digit permutation
// starting at kHaltonPerm[offset] (`ScrambledRadicalInverse`,
// core/volume.cu:1060-1076).  Host only: the 1024 x 4 sample table is built
// once per call on the host and shared by all texels.
static inline float scrambled_radical_inverse(unsigned base, unsigned offset, unsigned a) {
    const float inv_base = (float)1 / (float)base;
    unsigned reversed = 0;
    float inv_base_n = 1;
    while (a) {
        const unsigned next = a / base;
        const unsigned digit = a - next * base;
        reversed = reversed * base + kHaltonPerm[offset + digit];
        inv_base_n *= inv_base;
        a = next;
    }
    const float v = inv_base_n * ((float)reversed + inv_base * (float)kHaltonPerm[offset] / (1 - inv_base));
    return v < 0.99999994f ? v : 0.99999994f;
}

// table[4*k + c] for sample k: c = 0,1,2 -> Halton dims 0,1,2 (bases 2,3,5),
// c = 3 -> dim 4 (base 11).
static inline void halton_table(float* table) {
    for (unsigned k = 0; k < (unsigned)kHgLutSamples; ++k) {
        table[4 * k + 0] = scrambled_radical_inverse(2, 0, k);
        table[4 * k + 1] = scrambled_radical_inverse(3, 2, k);
        table[4 * k + 2] = scrambled_radical_inverse(5, 5, k);
        table[4 * k + 3] = scrambled_radical_inverse(11, 17, k);
    }
}

// The g-dependent blend exponent of the MIS weight (core/volume.cu:1117-1118),
// evaluated on the host in double as the reference does.
static inline float hg_lut_blend(float g) {
    double a = g < 0 ? -(double)g : (double)g;
    if (a < 0.6) return 0.0f;
    if (a < 0.7) return (float)((a - 0.6) * 9);
    return (float)(pow((a - 0.7) / 0.3, 0.05) * 0.1 + 0.9);
}

MRP_HD float frac_(float f) { return f - floorf(f); }

// One texel.  `angle` = ((j+0.5)/512) * pi/3 in double, rounded to float;
// `bn` = the texel's uniform offset; `blend` = hg_lut_blend(g).
MRP_HD float hg_lut_texel(int i, float angle, float g, float blend, float bn, const float* halton) {
    const float costheta = (((float)i + 0.5f) / (float)kHgLutSize) * 2.0f - 1.0f;
    float s_a, c_a;
    sincosf_(angle, &s_a, &c_a);
    const float radius = s_a / c_a;
    const vec3 view = v3(1.0f, 0.0f, 0.0f);
    const vec3 dir = v3(costheta, sqrtf(1.0f - costheta * costheta), 0.0f);
    float mis = 0.0f;
    {
        const float p = (radius * radius - 1.0f) + dir.x * dir.x;
        if (p > 0.0f) {
            const float sdotl = g > 0.0f ? dir.x : -dir.x;
            const float w = sqrtf(p);
            const float d0 = fmax_(0.0f, sdotl - w) / radius;
            const float d1 = (sdotl + w) / radius;
            const float d = d1 * d1 * d1 - d0 * d0 * d0;
            mis = (d1 > d0) ? d : 0.0f;
        }
        mis = fmin_(fmax_(mis, 1e-8f), 1.0f);
        mis = 1.0f - 1.0f / (1.0f + blend * mis / (1.0f - blend));
    }
    float hg = 0.0f;
    for (int k = 0; k < kHgLutSamples; k++) {
        const float h0 = halton[4 * k + 0], h1 = halton[4 * k + 1], h2 = halton[4 * k + 2], h4 = halton[4 * k + 3];
        if (frac_(h4 + bn) < mis) {
            const float rx = frac_(h0 + bn), ry = frac_(h1 + bn);
            const vec3 l = hg_sample_dir(rx, ry, view, g);
            const float sdotl = dot(dir, l);
            const float p = (radius * radius - 1.0f) + sdotl * sdotl;
            if (p <= 0.0f) continue;
            const float w = sqrtf(p);
            float d0 = sdotl - w;
            const float d1 = (sdotl + w) / radius;
            d0 = fmax_(0.0f, d0) / radius;
            const float d = d1 * d1 * d1 - d0 * d0 * d0;
            hg += fmax_(0.0f, d / 4.0f);
        } else {
            const float rx = frac_(h0 + bn), ry = frac_(h1 + bn), rz = frac_(h2 + bn);
            const float theta = rx * 6.28318548f;
            const float phi = acosf_(2.0f * ry - 1.0f);
            const float r = rz > 0.0f ? expf_(logf_(rz) * 0.333333343f) : 0.0f;  // cube root
            float st, ct, sp, cp;
            sincosf_(theta, &st, &ct);
            sincosf_(phi, &sp, &cp);
            const float x = radius * r * sp * ct;
            const float y = radius * r * sp * st;
            const float z = radius * r * cp;
            const vec3 rp = v3(x, y, z) + dir;
            const float c = dot(normalize(rp), view);
            hg += hg_phase(c, g);
        }
    }
    return hg / (float)kHgLutSamples;
}

}  // namespace mrp
