// mrp_frame.h - progressive-frame arithmetic shared by the gfx950 kernels and
// the CPU oracle: per-pixel histograms, the histogram-distance filter, tone
// mapping and the BGRA8 pack of `VolumeRender::Render(target, histo, target2,
// ...)` (core/volume.cu:57-124,181-264,469-547; core/render.cuh:640-732).
//
// IEEE-only like mrp_math.h: the same source compiled for gfx950 and for the
// host gives the same bits, so frames produced on the GPU can be compared
// with the oracle's exactly.
#pragma once

#include "../../include/mrp.h"
#include "mrp_medium.h"

namespace mrp {

constexpr int kHistoBins = MRP_HISTO_SIZE;  // bins per channel
constexpr int kHistoFloats = 3 * kHistoBins + 3;

// x^y for x >= 0 (the tone curves' pow) from the shared log and exp.
MRP_HD float powf_(float x, float y) {
    if (x != x) return x;
    if (x > 0.0f) {
        if (x == f_inf()) return y > 0.0f ? f_inf() : (y == 0.0f ? 1.0f : 0.0f);
        return expf_(y * logf_(x));
    }
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : f_inf());
    return f_nan();
}

// `lerp` as the reference writes it for colours (core/vector.cu:29-30).
MRP_HD float lerp_ref(float a, float b, float v) { return a * (1.0f - v) + b * v; }
MRP_HD vec3 lerp_ref(vec3 a, vec3 b, float v) {
    return v3(lerp_ref(a.x, b.x, v), lerp_ref(a.y, b.y, v), lerp_ref(a.z, b.z, v));
}

MRP_HD float saturate_(float v) { return fmin_(fmax_(v, 0.0f), 1.0f); }

// Row-major 3x3 times vector, each row a dot product (`float3x3::operator*`,
// core/vector.cu:256-258).
MRP_HD vec3 mat_vec(float a0, float a1, float a2, float b0, float b1, float b2, float c0, float c1, float c2, vec3 v) {
    return v3(a0 * v.x + a1 * v.y + a2 * v.z, b0 * v.x + b1 * v.y + b2 * v.z, c0 * v.x + c1 * v.y + c2 * v.z);
}

// `Gamma` (core/render.cuh:640-642): per channel x^(1/2.2).
MRP_HD vec3 tone_gamma(vec3 c) {
    const float e = 1.0f / 2.2f;
    return v3(powf_(c.x, e), powf_(c.y, e), powf_(c.z, e));
}

// `ACES` (core/render.cuh:643-732): sRGB -> AP0 -> AP1, desaturate, the
// RRT+ODT rational fit, dark-to-dim surround, desaturate, AP1 -> XYZ, D60 -> D65,
// XYZ -> Rec.709, then Gamma.  The matrices are the published ACES/sRGB ones.
MRP_HD vec3 tone_aces(vec3 color) {
    const vec3 y_of = v3(0.272229f, 0.674082f, 0.0536895f);
    vec3 aces = mat_vec(0.4397010f, 0.3829780f, 0.1773350f, 0.0897923f, 0.8134230f, 0.0967616f, 0.0175440f,
                        0.1115440f, 0.8707040f, color);
    vec3 acescg = mat_vec(1.4514393161f, -0.2365107469f, -0.2149285693f, -0.0765537734f, 1.1762296998f,
                          -0.0996759264f, 0.0083161484f, -0.0060324498f, 0.9977163014f, aces);
    float tmp = dot(acescg, y_of);
    acescg = lerp_ref(v3(tmp, tmp, tmp), acescg, 0.96f);
    const float a = 278.5085f, b = 10.7772f, c = 293.6045f, d = 88.7122f, e = 80.6889f;
    const vec3 x = acescg;
    const vec3 post = v3((x.x * (x.x * a + b)) / (x.x * (x.x * c + d) + e),
                         (x.y * (x.y * a + b)) / (x.y * (x.y * c + d) + e),
                         (x.z * (x.z * a + b)) / (x.z * (x.z * c + d) + e));
    // darkSurround_to_dimSurround
    vec3 xyz = mat_vec(0.6624541811f, 0.1340042065f, 0.1561876870f, 0.2722287168f, 0.6740817658f, 0.0536895174f,
                       -0.0055746495f, 0.0040607335f, 1.0103391003f, post);
    const float divisor = fmax_(xyz.x + xyz.y + xyz.z, 1e-4f);
    vec3 xyy = v3(xyz.x / divisor, xyz.y / divisor, xyz.y);
    xyy.z = fmin_(fmax_(xyy.z, 0.0f), 65504.0f);
    xyy.z = powf_(xyy.z, 0.9811f);
    const float m = xyy.z / fmax_(xyy.y, 1e-4f);
    xyz = v3(xyy.x, xyy.z, (1.0f - xyy.x - xyy.y));
    xyz.x = xyz.x * m;
    xyz.z = xyz.z * m;
    vec3 lin = mat_vec(1.6410233797f, -0.3248032942f, -0.2364246952f, -0.6636628587f, 1.6153315917f, 0.0167563477f,
                       0.0117218943f, -0.0082844420f, 0.9883948585f, xyz);
    tmp = dot(lin, y_of);
    lin = lerp_ref(v3(tmp, tmp, tmp), lin, 0.93f);
    xyz = mat_vec(0.6624541811f, 0.1340042065f, 0.1561876870f, 0.2722287168f, 0.6740817658f, 0.0536895174f,
                  -0.0055746495f, 0.0040607335f, 1.0103391003f, lin);
    xyz = mat_vec(0.98722400f, -0.00611327f, 0.0159533f, -0.00759836f, 1.00186000f, 0.0053302f, 0.00307257f,
                  -0.00509595f, 1.0816800f, xyz);
    lin = mat_vec(3.2409699419f, -1.5373831776f, -0.4986107603f, -0.9692436363f, 1.8759675015f, 0.0415550574f,
                  0.0556300797f, -0.2039769589f, 1.0569715142f, xyz);
    return tone_gamma(lin);
}

// Exposure, tone curve (0 none, 1 gamma, 2 ACES) and the 0xAARRGGBB pack of
// `Denoise` / `ReprojectionDenoise` (core/volume.cu:248-263,276-286).
MRP_HD uint32_t tone_pack(vec3 res, float exposure, int tone) {
    vec3 val = res * exposure;
    if (tone == 1) val = tone_gamma(val);
    else if (tone == 2) val = tone_aces(val);
    const uint32_t r = (uint32_t)(255.0f * saturate_(val.x));
    const uint32_t g = (uint32_t)(255.0f * saturate_(val.y));
    const uint32_t b = (uint32_t)(255.0f * saturate_(val.z));
    return 0xff000000u | (r << 16) | (g << 8) | b;
}

// Frame index used for the running means: with the checkerboard on, each
// pixel is rendered every second call (core/volume.cu:101).
MRP_HD int frame_index(int frame, int checkerboard) { return checkerboard ? frame / 2 : frame; }

// Path-tracer branch of `RenderCamera` (core/volume.cu:103-123): running mean
// of the clamped radiance and the pixel's histogram of r/(r+1).
MRP_HD void histo_accumulate(float* target3, float* h /* kHistoFloats */, vec3 res, int fnum) {
    if (fnum == 0)
        for (int i = 0; i < kHistoFloats; ++i) h[i] = 0.0f;
    const float rate = 1.0f / (float)(1 + fnum);
    const vec3 t = lerp(v3(target3[0], target3[1], target3[2]), res, rate);
    target3[0] = t.x;
    target3[1] = t.y;
    target3[2] = t.z;
    const vec3 r = v3(res.x / (res.x + 1.0f), res.y / (res.y + 1.0f), res.z / (res.z + 1.0f));
    h[3 * kHistoBins] += 1.0f;  // total sample count
    const int bx = (int)(fmin_(r.x, 0.999f) * (float)kHistoBins);
    const int by = (int)(fmin_(r.y, 0.999f) * (float)kHistoBins);
    const int bz = (int)(fmin_(r.z, 0.999f) * (float)kHistoBins);
    h[bx] += 1.0f;
    h[by + kHistoBins] += 1.0f;
    h[bz + 2 * kHistoBins] += 1.0f;
    const float l = r.x + r.y + r.z;
    h[3 * kHistoBins + 2] = lerp(h[3 * kHistoBins + 2], l, rate);      // x
    h[3 * kHistoBins + 1] = lerp(h[3 * kHistoBins + 1], l * l, rate);  // x2
}

// Chi-square-like distance of two pixel histograms (`Compare`,
// core/volume.cu:187-207): sum over non-empty bins of
// (sqrt(ny/nx) hx - sqrt(nx/ny) hy)^2 / (hx + hy), divided by the bin count.
MRP_HD float histo_distance(const float* hx_bins, float nx, const float* hy_bins, float ny) {
    const float syx = sqrtf(ny / nx), sxy = sqrtf(nx / ny);
    float p = 0.0f, res = 0.0f;
    for (int i = 0; i < 3 * kHistoBins; ++i) {
        const float hx = hx_bins[i], hy = hy_bins[i];
        if (hx != 0.0f || hy != 0.0f) {
            p += 1.0f;
            const float t = syx * hx - sxy * hy;
            res += t * t / (hx + hy);
        }
    }
    return p == 0.0f ? 0.0f : res / p;
}

// Filter weight of one neighbour: max(0, 1 - 1.2 d), evaluated in double as the
// reference's double literal makes it (core/volume.cu:236).
MRP_HD float histo_weight(float dist) {
    const double w = 1.0 - 1.2 * (double)dist;
    return (float)(w > 0.0 ? w : 0.0);
}

// Clamp a neighbour coordinate into the image (`UnRoll`, core/volume.cu:181-185).
MRP_HD int clampi_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Reprojection of the MRPNN branch of `RenderCamera` (core/volume.cu:125-141):
// the pixel that saw `motion_pos` from the previous camera.  The reference
// assumes unit right/up vectors (a 90-degree camera); dividing by their
// squared lengths keeps the mapping right for any field of view and changes
// nothing for unit vectors.
MRP_HD int reproject_pixel(vec3 motion_pos, vec3 lori, vec3 lup, vec3 lright, int w, int h) {
    vec3 ldir = motion_pos - lori;
    const vec3 lforward = normalize(lori);
    ldir = ldir / dot(ldir, lforward);
    ldir = ldir - lforward;
    const float lu = dot(ldir, lright) / dot(lright, lright) * 0.5f + 0.5f;
    const float lv = dot(ldir, lup) / dot(lup, lup) * -0.5f + 0.5f;
    const float fu = lu * (float)w, fv = lv * (float)h;
    // int() of a float outside the int range is undefined: clamp in float first
    const int x = fu != fu ? 0 : (int)fmin_(fmax_(fu, -1.0f), (float)w);
    const int y = fv != fv ? 0 : (int)fmin_(fmax_(fv, -1.0f), (float)h);
    return clampi_(y, 0, h - 1) * w + clampi_(x, 0, w - 1);
}

}  // namespace mrp
