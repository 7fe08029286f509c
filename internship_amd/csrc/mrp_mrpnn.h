// mrp_mrpnn.h - constants of the MRPNN descriptor shared by the kernels and the oracle.
//
// The descriptor of a scattering point has 192 taps.  Taps 0..159 sit on eight
// shells around the point (stencil points of the 8 mip levels, oriented by the
// random frame X,Y,Z of the view direction); taps 160..191 march along the
// light direction over mips 0..3.  Per tap the schedule fixes an offset (box
// units), a density mip and the v-coordinate of the cone-averaged HG table.
// Follows `GetSamples23_` (core/vector.cu:299-400) and the per-tap loop of
// `RadiancePredict` (core/radiancePredict.cu:321-359).
#pragma once

#include "mrp_math.h"
#include "mrp_tables_gen.h"

namespace mrp {

constexpr int kTaps = 192;         // descriptor taps
constexpr int kShellTaps = 160;    // taps placed on stencil shells
constexpr int kFeatures = 3 * kTaps;

struct TapTable {
    float offset[kTaps];
    int mip[kTaps];
    float lut_v[kTaps];            // HG-table v of taps 1..191 (tap 0 uses HG directly)
    float stencil[kShellTaps][3];  // unit-frame stencil points
    // HG-table u of the shell taps.  A shell tap lies at p + (X s0 + Y s1 + Z s2) * off
    // with (X, Y, Z) orthonormal and off > 0, so its direction from p is
    // (X s0 + Y s1 + Z s2) / |s|: its cosine to X is s0 / |s| for every vertex, and
    // its cosine to the light axis LX is (LX.X s0 + LX.Y s1 + LX.Z s2) / |s|.
    float inv_len[kShellTaps];     // 1 / |s|
    float u_view[kShellTaps];      // s0 / |s| * 0.5 + 0.5
};

// reference lerp(a, b, v) = a*(1-v) + b*v in FP32
static inline float sched_lerp(float a, float b, float v) { return a * (1.0f - v) + b * v; }

static inline void descriptor_schedule(TapTable* t) {
    static const float kStencil[kShellTaps][3] = MRP_STENCIL_INIT;
    const float ms = 0.6f;  // MsOffsetScale
    for (int i = 0; i < kTaps; ++i) {
        float off;
        int mip;
        if (i < 8) {
            mip = 0;
            off = (float)(1.0 / 256.0);
            off *= sched_lerp(ms, 1.0f, 1.0f);
        } else if (i < 16) {
            mip = 1;
            off = (float)(2.5 / 256.0);
            off *= sched_lerp(ms, 1.0f, 0.8f);
        } else if (i < 32) {
            mip = 2;
            off = (float)(5.5 / 256.0);
            off *= sched_lerp(ms, 1.0f, 0.6f);
        } else if (i < 48) {
            mip = 3;
            off = (float)(11.5 / 256.0);
            off *= sched_lerp(ms, 1.0f, 0.4f);
        } else if (i < 64) {
            mip = 4;
            off = (float)(23.5 / 256.0);
            off *= sched_lerp(ms, 1.0f, 0.2f);
        } else if (i < 96) {
            mip = 5;
            off = (float)(47.5 / 256.0);
            off *= ms;
        } else if (i < 128) {
            mip = 6;
            off = (float)(95.5 / 256.0);
            off *= ms;
        } else if (i < 160) {
            mip = 7;
            off = (float)(191.5 / 256.0);
            off *= ms;
        } else if (i < 168) {
            mip = 0;
            off = (float)(i - 160 + 1);
            off = (float)((double)off / 256.0);
        } else if (i < 176) {
            mip = 1;
            off = (float)(8.5 + 1 + 2.0 * (i - 168));
            off = (float)((double)off / 256.0);
        } else if (i < 184) {
            mip = 2;
            off = (float)(26.5 + 4.0 * (i - 176));
            off = (float)((double)off / 256.0);
        } else {
            mip = 3;
            off = (float)(60.5 + 8.0 * (i - 184));
            off = (float)((double)off / 256.0);
        }
        t->offset[i] = off;
        t->mip[i] = mip;
        const float radius = (float)(1 << mip) / 256.0f;
        const float angle = atanf_(0.5f * radius / off);
        t->lut_v[i] = angle / (3.1415926535f * 60.0f / 180.0f);
    }
    for (int i = 0; i < kShellTaps; ++i) {
        for (int c = 0; c < 3; ++c) t->stencil[i][c] = kStencil[i][c];
        const float* s = kStencil[i];
        t->inv_len[i] = 1.0f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        t->u_view[i] = s[0] * t->inv_len[i] * 0.5f + 0.5f;
    }
}

// HG-table u of a shell tap's direction against the light axis (see TapTable),
// from lxX = LX.X, lxY = LX.Y, lxZ = LX.Z.
MRP_HD float tap_u_light(float lxX, float lxY, float lxZ, float s0, float s1, float s2, float inv_len) {
    return ((lxX * s0 + lxY * s1) + lxZ * s2) * inv_len * 0.5f + 0.5f;
}

}  // namespace mrp
