// mrp_volume_render.hpp - C++ façade with the reference's class interface over
// the C ABI of libmrp (mrp.h).  Header-only; no HIP or torch types.
//
// It keeps the signatures of `VolumeRender` (core/volume.hpp:45-146) and
// `Camera` (core/camera.hpp:10-41), so that code written against the
// reference (main.cu:9-47, the README example) compiles against this header
// with `using namespace internship;`.  The differences:
//   * `float3`, `int2` and `Histogram` are this header's PODs (same layouts as
//     CUDA's float3/int2 and the reference's Histogram);
//   * errors print the library's message and continue, as the reference's
//     CheckError does (core/volume.cu:15); LastError() returns it;
//   * without SetHDRI the environment is a constant sky colour (SetSky;
//     SetEnvExp scales it, or the map once one is set);
//   * MRPNN rendering needs the trained weights: LoadNetwork(path) reads the
//     49,673-float blob (internship_amd/data/mrpnn_weights.bin), or
//     SetNetwork(weights, count) takes them from memory.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "mrp.h"

namespace internship {

struct float3 {
    float x, y, z;
};
struct int2 {
    int x, y;
};
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline float3 operator*(float3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float3 operator-(float3 a) { return {-a.x, -a.y, -a.z}; }
inline float3 operator+(float3 a, float3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator+(float3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }

typedef mrp_histogram Histogram;
typedef std::function<float(int x, int y, int z, float u, float v, float w)> FillFunc;  // omp.hpp:10

namespace detail {
inline mrp_vec3 v(float3 a) { return mrp_vec3{a.x, a.y, a.z}; }
inline float3 normalize(float3 a) {
    const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}
inline float3 cross(float3 a, float3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline bool ok(int rc, const char* what) {
    if (rc != MRP_OK) std::printf("%s: %s\n", what, mrp_last_error());
    return rc == MRP_OK;
}
inline float lerp(float a, float b, float t) { return a + (b - a) * t; }
// `Sample` / `SampleClamp` (core/volume.cu:675-722): FP32 trilinear at
// uv * res - 0.5 with border-0 (negative values read as 0) or clamp-to-edge corners.
inline float sample(const std::vector<float>& d, int res, float3 uv, bool clamp) {
    auto at = [&](int x, int y, int z) -> float {
        if (clamp) {
            x = std::max(std::min(x, res - 1), 0);
            y = std::max(std::min(y, res - 1), 0);
            z = std::max(std::min(z, res - 1), 0);
        } else if (x < 0 || y < 0 || z < 0 || x >= res || y >= res || z >= res) {
            return 0.0f;
        }
        return std::max(0.0f, d[((size_t)x * res + y) * res + z]);
    };
    const float px = uv.x * res - 0.5f, py = uv.y * res - 0.5f, pz = uv.z * res - 0.5f;
    const int x = (int)std::floor(px), y = (int)std::floor(py), z = (int)std::floor(pz);
    const float wx = px - x, wy = py - y, wz = pz - z;
    return lerp(lerp(lerp(at(x, y, z), at(x, y, z + 1), wz), lerp(at(x, y + 1, z), at(x, y + 1, z + 1), wz), wy),
                lerp(lerp(at(x + 1, y, z), at(x + 1, y, z + 1), wz),
                     lerp(at(x + 1, y + 1, z), at(x + 1, y + 1, z + 1), wz), wy),
                wx);
}
// `RayBoxOffset` (core/vector.cu:41-58): entry distance into the 0.4999 slab box, -1 on a miss.
inline float ray_box_offset(float3 p, float3 d) {
    const float id[3] = {d.x == 0 ? 999999999.0f : 1.0f / d.x, d.y == 0 ? 999999999.0f : 1.0f / d.y,
                         d.z == 0 ? 999999999.0f : 1.0f / d.z};
    const float pp[3] = {p.x, p.y, p.z};
    float tmi = -1e30f, tma = 1e30f;
    for (int k = 0; k < 3; ++k) {
        const float a = -pp[k] * id[k], b = 0.4999f * id[k];
        tmi = std::max(tmi, std::min(a + b, a - b));
        tma = std::min(tma, std::max(a + b, a - b));
    }
    return tma >= tmi ? std::max(tmi, 0.0f) : -1.0f;
}
// HBM staging for the point queries (the reference allocates per call too).
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { ok(mrp_malloc(&p, bytes ? bytes : 16), "mrp_malloc"); }
    ~DevBuf() { mrp_free(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    float* f() const { return (float*)p; }
};
}  // namespace detail

// The volume files of `VolumeRender(string path)` (core/volume.cu:835-911):
// `path.bin` (int32 resolution + FP32 grid) if present, else `path` as `.txt`
// (the resolution, then the values) or `.vox` (256^3 byte values as text, 64
// per line after w/h/d header lines, x 64/255).  No GPU work: the constructor
// uploads what this reads.
inline bool LoadVolumeFile(const std::string& path, int* resolution, std::vector<float>* datas, bool* from_bin) {
    *from_bin = false;
    if (FILE* f = std::fopen((path + ".bin").c_str(), "rb")) {
        bool good = std::fread(resolution, sizeof(int), 1, f) == 1 && *resolution > 0 && *resolution <= 1024;
        if (good) {
            datas->assign((size_t)*resolution * *resolution * *resolution, 0.0f);
            good = std::fread(datas->data(), sizeof(float), datas->size(), f) == datas->size();
        }
        std::fclose(f);
        *from_bin = good;
        return good;
    }
    const std::string fmt = path.size() >= 3 ? path.substr(path.size() - 3) : "";
    if (fmt == "txt") {
        std::ifstream f(path);
        if (!(f >> *resolution) || *resolution <= 0 || *resolution > 1024) return false;
        datas->assign((size_t)*resolution * *resolution * *resolution, 0.0f);
        for (float& d : *datas)
            if (!(f >> d)) break;
        return true;
    }
    if (fmt == "vox") {
        *resolution = 256;
        datas->assign((size_t)256 * 256 * 256, 0.0f);
        std::ifstream f(path);
        std::string line;
        size_t i = 0;
        const float inv = (float)(64.0 / 255.0);
        while (std::getline(f, line) && i < datas->size()) {
            if (line.empty() || line[0] == 'w' || line[0] == 'h' || line[0] == 'd') continue;
            std::stringstream s(line);
            int d;
            for (int j = 0; j < 64 && (s >> d) && i < datas->size(); ++j) (*datas)[i++] = (float)d * inv;
        }
        return true;
    }
    return false;
}

class VolumeRender {
   public:
    enum RenderType { PT, RPNN, MRPNN };

    int resolution = 0;
    float max_density = 0.00001f;

    explicit VolumeRender(int res) : resolution(res), datas_((size_t)res * res * res, 0.0f) {
        detail::ok(mrp_volume_create(res, &h_), "VolumeRender");
    }

    // `VolumeRender(string path)` (core/volume.cu:835-911): `path.bin` (int32
    // resolution + FP32 grid) if present, else `.txt` (resolution, then the
    // values) or `.vox` (256^3 bytes as text, x 64/255); writes `path.bin`.
    explicit VolumeRender(const std::string& path) {
        if (!LoadVolumeFile(path, &resolution, &datas_, &have_bin_)) {
            std::printf("File not found!\n");
            return;
        }
        detail::ok(mrp_volume_create(resolution, &h_), "VolumeRender");
        Update();
        if (!have_bin_) {
            if (FILE* f = std::fopen((path + ".bin").c_str(), "wb")) {
                std::fwrite(&resolution, sizeof(int), 1, f);
                std::fwrite(datas_.data(), sizeof(float), datas_.size(), f);
                std::fclose(f);
            }
        }
    }

    VolumeRender(const VolumeRender&) = delete;
    VolumeRender& operator=(const VolumeRender&) = delete;

    ~VolumeRender() {
        if (net_) mrp_network_destroy(net_);
        if (h_) mrp_volume_destroy(h_);
    }

    void SetData(int x, int y, int z, float value) { datas_[((size_t)x * resolution + y) * resolution + z] = value; }

    // `ParallelFill` (core/omp.cpp:4-18): f(x, y, z, u, v, w) at voxel centres.
    void SetDatas(FillFunc f) {
        const int r = resolution;
#pragma omp parallel for
        for (int i = 0; i < r; i++)
            for (int j = 0; j < r; j++)
                for (int k = 0; k < r; k++)
                    datas_[((size_t)i * r + j) * r + k] =
                        f(i, j, k, (float)((i + 0.5) / r), (float)((j + 0.5) / r), (float)((k + 0.5) / r));
    }

    void Update() {
        if (!h_) return;
        host_mips_.clear();
        host_trs_.clear();
        host_lut_.clear();
        if (!detail::ok(mrp_volume_set_data(h_, datas_.data(), 0), "Update")) return;
        if (!detail::ok(mrp_volume_update(h_), "Update")) return;
        detail::ok(mrp_volume_max_density(h_, &max_density), "Update");
    }
    void Update_TR(float3 lightDir, float alpha = 64.0f, bool /*CPU*/ = false) {
        host_trs_.clear();
        detail::ok(mrp_volume_update_tr(h_, detail::v(lightDir), alpha), "Update_TR");
    }
    void UpdateHGLut(float g) {
        host_lut_.clear();
        detail::ok(mrp_volume_update_hg_lut(h_, g, seed_), "UpdateHGLut");
    }

    // ---- host queries (core/volume.cu:1184-1201,1495-1531): on host copies of
    // the density mips, transmittance mips and HG table, fetched on first use
    // after Update / Update_TR / UpdateHGLut (and after MRPNN renders and
    // queries, which refresh the light's tables themselves).
    float DensityAtUV(int mip, float3 uv) {
        if (mip < 0 || mip > 8 || !fetch_mips()) return 0.0f;
        return detail::sample(host_mips_[mip], 256 >> mip, uv, false);
    }
    float DensityAtUV(float mip, float3 uv) {
        const int a = (int)mip;
        return detail::lerp(DensityAtUV(a, uv), DensityAtUV(a + 1, uv), mip - a);
    }
    float DensityAtPosition(int mip, float3 pos) { return DensityAtUV(mip, pos + 0.5f); }
    float DensityAtPosition(float mip, float3 pos) {
        const int a = (int)mip;
        return detail::lerp(DensityAtPosition(a, pos), DensityAtPosition(a + 1, pos), mip - a);
    }
    // A point outside the box reads the transmittance where its light ray enters
    // the box, 1 if the ray misses it (`TrAtPosition`, core/volume.cu:1498-1515).
    float TrAtPosition(int mip, float3 pos, float3 lightDir) {
        if (mip < 0 || mip > 7 || !fetch_trs()) return 0.0f;
        const int r = 128 >> mip;
        if (pos.x < -0.5f || pos.y < -0.5f || pos.z < -0.5f || pos.x > 0.5f || pos.y > 0.5f || pos.z > 0.5f) {
            const float off = detail::ray_box_offset(pos, lightDir);
            if (off < 0.0f) return 1.0f;
            return detail::sample(host_trs_[mip], r, pos + 0.5f + lightDir * off, true);
        }
        return detail::sample(host_trs_[mip], r, pos + 0.5f, true);
    }
    // The cone-averaged HG table at (cos, cone angle in radians, at most 60
    // degrees): bilinear on the texel corners u = cos/2 + 1/2, v = angle/60deg as
    // the reference indexes it (core/volume.cu:1184-1201); each axis is weighted
    // by its own fraction (the reference weights u by v's fraction and v by u's).
    float GetHGLut(float cos, float angle) {
        if (!fetch_lut()) return 0.0f;
        const int N = 512;
        const float u = cos * 0.5f + 0.5f, v = angle / (3.1415926535f * 60.0f / 180.0f);
        auto idx = [&](float t, int k) { return std::max(std::min((int)(t * N) + k, N - 1), 0); };
        const int u0 = idx(u, 0), u1 = idx(u, 1), v0 = idx(v, 0), v1 = idx(v, 1);
        const float fu = u * N - std::floor(u * N), fv = v * N - std::floor(v * N);
        const float a0 = detail::lerp(host_lut_[u0 + v0 * N], host_lut_[u1 + v0 * N], fu);
        const float a1 = detail::lerp(host_lut_[u0 + v1 * N], host_lut_[u1 + v1 * N], fu);
        return detail::lerp(a0, a1, fv);
    }

    // ---- GPU point queries
    // `GetTr` (core/volume.cu:302-323): the mean of Tr(ori -> light) * HG(dir .
    // light) over sampleNum estimates rounded up to a multiple of 32.
    float3 GetTr(float3 ori, float3 dir, float3 lightDir, float alpha = 1, float g = 0, int sampleNum = 1) {
        const int spp = (std::max(sampleNum, 1) + 31) / 32 * 32;
        detail::DevBuf po(12), pd(12), out(4);
        float r = 0.0f;
        if (detail::ok(mrp_memcpy_h2d(po.p, &ori, 12), "GetTr") && detail::ok(mrp_memcpy_h2d(pd.p, &dir, 12), "GetTr") &&
            detail::ok(mrp_get_shadow_terms(h_, 1, po.f(), pd.f(), detail::v(lightDir), alpha, g, spp, seed_, out.f(),
                                            nullptr), "GetTr"))
            detail::ok(mrp_memcpy_d2h(&r, out.p, 4), "GetTr");
        return {r, r, r};
    }
    // `GetTrs` (core/volume.cu:636-673): lightColor * pi * the mean of Tr * HG per point.
    std::vector<float3> GetTrs(float alpha, const std::vector<float3>& ori, const std::vector<float3>& dir,
                               float3 lightDir, float3 lightColor, float g = 0, int sampleNum = 1) {
        const size_t n = std::min(ori.size(), dir.size());
        std::vector<float> t(n, 0.0f);
        std::vector<float3> out(n, float3{0, 0, 0});
        if (!n) return out;
        detail::DevBuf po(n * 12), pd(n * 12), pt(n * 4);
        if (detail::ok(mrp_memcpy_h2d(po.p, ori.data(), n * 12), "GetTrs") &&
            detail::ok(mrp_memcpy_h2d(pd.p, dir.data(), n * 12), "GetTrs") &&
            detail::ok(mrp_get_shadow_terms(h_, (int)n, po.f(), pd.f(), detail::v(lightDir), alpha, g,
                                            std::max(sampleNum, 1), seed_, pt.f(), nullptr), "GetTrs") &&
            detail::ok(mrp_memcpy_d2h(t.data(), pt.p, n * 4), "GetTrs"))
            for (size_t i = 0; i < n; ++i) {
                const float s = t[i] * 3.1415926535f;
                out[i] = {lightColor.x * s, lightColor.y * s, lightColor.z * s};
            }
        return out;
    }
    // `GetRadiances` (core/volume.cu:329-402): PT - the mean of sampleNum paths
    // per ray; MRPNN - one `NNPredict` per ray (the light's tables are refreshed
    // as the reference's UpdateHGLut / Update_TR do).
    std::vector<float3> GetRadiances(const std::vector<float3>& ori, const std::vector<float3>& dir, float3 lightDir,
                                     float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1, float g = 0,
                                     int sampleNum = 1, RenderType rt = PT) {
        const size_t n = std::min(ori.size(), dir.size());
        std::vector<float3> out(n, float3{0, 0, 0});
        if (!n) return out;
        mrp_render_params p = params(int2{1, 1}, float3{0, 0, 0}, float3{0, 1, 0}, float3{1, 0, 0}, lightDir, rt, g,
                                     alpha, lightColor, multiScatter);
        p.spp_count = std::max(sampleNum, 1);
        std::vector<float> r(n * 4);
        detail::DevBuf po(n * 12), pd(n * 12), pr(n * 16);
        if (rt != PT) {
            host_trs_.clear();
            host_lut_.clear();
        }
        if (detail::ok(mrp_memcpy_h2d(po.p, ori.data(), n * 12), "GetRadiances") &&
            detail::ok(mrp_memcpy_h2d(pd.p, dir.data(), n * 12), "GetRadiances") &&
            detail::ok(mrp_get_radiances(h_, (int)n, po.f(), pd.f(), &p, pr.f(), nullptr), "GetRadiances") &&
            detail::ok(mrp_memcpy_d2h(r.data(), pr.p, n * 16), "GetRadiances"))
            for (size_t i = 0; i < n; ++i) out[i] = {r[4 * i], r[4 * i + 1], r[4 * i + 2]};
        return out;
    }
    // `GetSamples` (core/volume.cu:550-626): the training target of `GetSample`
    // at each point (high-order in-scattering, sampleNum paths of up to
    // multiScatter bounces), per-point extinction, light, g and scatter.
    std::vector<float3> GetSamples(const std::vector<float>& alpha, const std::vector<float3>& ori,
                                   const std::vector<float3>& dir, const std::vector<float3>& lightDir,
                                   const std::vector<float>& g, const std::vector<float>& scatter,
                                   float3 lightColor = {1, 1, 1}, int multiScatter = 1, int sampleNum = 1) {
        const size_t n = std::min({alpha.size(), ori.size(), dir.size(), lightDir.size(), g.size(), scatter.size()});
        std::vector<float3> out(n, float3{0, 0, 0});
        if (!n) return out;
        std::vector<mrp_sample_query> q(n);
        for (size_t i = 0; i < n; ++i)
            q[i] = mrp_sample_query{detail::v(ori[i]), detail::v(dir[i]), detail::v(lightDir[i]), alpha[i], g[i],
                                    scatter[i]};
        detail::DevBuf pq(n * sizeof(mrp_sample_query)), po(n * 12);
        if (detail::ok(mrp_memcpy_h2d(pq.p, q.data(), n * sizeof(mrp_sample_query)), "GetSamples") &&
            detail::ok(mrp_get_samples(h_, (int)n, (const mrp_sample_query*)pq.p, detail::v(lightColor),
                                       detail::v(sky_), multiScatter, std::max(sampleNum, 1), seed_, po.f(),
                                       nullptr), "GetSamples"))
            detail::ok(mrp_memcpy_d2h(out.data(), po.p, n * 12), "GetSamples");
        return out;
    }

    // setters of the reference's device globals (core/volume.cu:1460-1493)
    void SetCheckboard(bool on) { checkboard_ = on; }
    void SetEnvExp(float e) {  // scales the environment map, or the constant sky when none is set
        env_exp_ = e;
        sky_ = base_sky_ * e;
        if (hdri_w_ > 0)
            detail::ok(mrp_volume_set_hdri(h_, hdri_.data(), hdri_w_, hdri_h_, env_exp_), "SetEnvExp");
    }
    void SetTrScale(float s) { tr_scale_ = s; }
    void SetScatterRate(float r) { scatter_ = {r, r, r}; }
    void SetScatterRate(float3 r) { scatter_ = r; }
    void SetExposure(float e) { exposure_ = e; }
    // `SetHDRI` (core/volume.cu:1426-1458): a Radiance .hdr environment map;
    // MRPNN renders then also sample it as a light (mrp_volume_set_hdri).
    void SetHDRI(const std::string& path) {
        int w = 0, h = 0;
        if (!detail::ok(mrp_read_hdr(path.c_str(), nullptr, 0, &w, &h), "SetHDRI")) return;
        std::vector<float> px((size_t)w * h * 4);
        if (!detail::ok(mrp_read_hdr(path.c_str(), px.data(), (size_t)w * h, &w, &h), "SetHDRI")) return;
        if (!detail::ok(mrp_volume_set_hdri(h_, px.data(), w, h, env_exp_), "SetHDRI")) return;
        hdri_.swap(px);
        hdri_w_ = w;
        hdri_h_ = h;
    }
    // this build's extras
    void SetSky(float3 c) {
        base_sky_ = c;
        sky_ = c * env_exp_;
    }
    void SetSeed(uint64_t s) { seed_ = s; }
    void SetMlpPrecision(int p) { precision_ = p; }  // 0 FP32, 1 BF16 MFMA, 2 FP8-E4M3 MFMA
    bool SetNetwork(const float* w, size_t count) {
        if (net_) mrp_network_destroy(net_);
        net_ = nullptr;
        if (!detail::ok(mrp_network_create(w, count, &net_), "SetNetwork")) return false;
        return detail::ok(mrp_volume_set_network(h_, net_), "SetNetwork");
    }
    bool LoadNetwork(const std::string& path) {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) {
            std::printf("LoadNetwork: cannot open %s\n", path.c_str());
            return false;
        }
        std::fseek(f, 0, SEEK_END);
        const long bytes = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        std::vector<float> w(bytes > 0 ? (size_t)bytes / sizeof(float) : 0);
        const size_t got = std::fread(w.data(), sizeof(float), w.size(), f);
        std::fclose(f);
        return SetNetwork(w.data(), got);
    }
    const char* LastError() const { return mrp_last_error(); }
    mrp_volume* Handle() { return h_; }

    // Offline render (core/volume.cu:410-467): the mean of sampleNum samples.
    std::vector<float3> Render(int2 size, float3 ori, float3 up, float3 right, float3 lightDir, RenderType rt = PT,
                               float g = 0.857f, float alpha = 1, float3 lightColor = {1, 1, 1},
                               int multiScatter = 512, int sampleNum = 1024) {
        std::vector<float3> out((size_t)size.x * size.y, float3{0, 0, 0});
        if (rt != PT) {
            host_trs_.clear();
            host_lut_.clear();
        }
        mrp_render_params p = params(size, ori, up, right, lightDir, rt, g, alpha, lightColor, multiScatter);
        p.spp_begin = 0;
        p.spp_count = sampleNum;
        detail::ok(mrp_render_offline(h_, &p, reinterpret_cast<float*>(out.data()), nullptr), "Render");
        return out;
    }

    // Progressive frame (core/volume.cu:469-547); target, histo_buffer and
    // target2 are device pointers owned by the caller.
    void Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, float3 ori, float3 up,
                float3 right, float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1,
                float g = 0, int randseed = 0, RenderType rt = PT, int toneType = 2, bool denoise = false) {
        mrp_frame_params fp;
        fp.r = params(size, ori, up, right, lightDir, rt, g, alpha, lightColor, multiScatter);
        fp.r.spp_begin = randseed;
        fp.r.spp_count = 1;
        fp.tone = toneType;
        fp.denoise = denoise ? 1 : 0;
        fp.checkerboard = checkboard_ ? 1 : 0;
        fp.exposure = exposure_;
        detail::ok(mrp_render_frame(h_, &fp, reinterpret_cast<float*>(target), histo_buffer, target2, nullptr),
                   "Render");
    }

   private:
    bool fetch_mips() {
        if (!host_mips_.empty()) return true;
        std::vector<std::vector<float>> m(9);
        for (int k = 0; k < 9; ++k) {
            const size_t r = (size_t)(256 >> k);
            m[k].resize(r * r * r);
            if (!detail::ok(mrp_volume_get_mip(h_, k, m[k].data()), "DensityAtUV")) return false;
        }
        host_mips_.swap(m);
        return true;
    }
    bool fetch_trs() {
        if (!host_trs_.empty()) return true;
        std::vector<std::vector<float>> m(8);
        for (int k = 0; k < 8; ++k) {
            const size_t r = (size_t)(128 >> k);
            m[k].resize(r * r * r);
            if (!detail::ok(mrp_volume_get_tr_mip(h_, k, m[k].data()), "TrAtPosition")) return false;
        }
        host_trs_.swap(m);
        return true;
    }
    bool fetch_lut() {
        if (!host_lut_.empty()) return true;
        std::vector<float> t((size_t)512 * 512);
        if (!detail::ok(mrp_volume_get_hg_lut(h_, t.data()), "GetHGLut")) return false;
        host_lut_.swap(t);
        return true;
    }

    mrp_render_params params(int2 size, float3 ori, float3 up, float3 right, float3 lightDir, RenderType rt, float g,
                             float alpha, float3 lightColor, int multiScatter) const {
        mrp_render_params p;
        p.width = size.x;
        p.height = size.y;
        p.ori = detail::v(ori);
        p.up = detail::v(up);
        p.right = detail::v(right);
        p.light_dir = detail::v(lightDir);
        p.light_color = detail::v(lightColor);
        p.sky = detail::v(sky_);
        p.scatter_rate = detail::v(scatter_);
        p.g = g;
        p.alpha = alpha;
        p.tr_scale = tr_scale_;
        p.multi_scatter = multiScatter;
        p.spp_begin = 0;
        p.spp_count = 1;
        p.seed = seed_;
        p.render_type = rt == PT ? MRP_RENDER_PT : MRP_RENDER_MRPNN;
        p.mlp_precision = precision_;
        return p;
    }

    mrp_volume* h_ = nullptr;
    mrp_network* net_ = nullptr;
    std::vector<float> datas_;
    bool have_bin_ = false;
    bool checkboard_ = true;
    float3 base_sky_{0.1f, 0.1f, 0.1f}, sky_{0.1f, 0.1f, 0.1f};
    float env_exp_ = 1.0f;
    float3 scatter_{1.001f, 1.001f, 1.001f};  // core/render.cuh:16
    float tr_scale_ = 1.0f;
    float exposure_ = 1.0f;
    uint64_t seed_ = 0;
    int precision_ = 1;
    std::vector<std::vector<float>> host_mips_, host_trs_;  // host copies for the queries
    std::vector<float> hdri_;  // the environment map's texels (re-sent when the exposure changes)
    int hdri_w_ = 0, hdri_h_ = 0;
    std::vector<float> host_lut_;
};

typedef std::function<float3(float3 radiance)> ToneMapFunction;

// `SaveBMP` (core/camera.cu:10-64): 32-bit BGRA, top-down; writes `<filename>.bmp`.
inline int SaveBMP(unsigned char* image, int w, int h, const char* filename) {
    unsigned char header[54] = {0x42, 0x4d, 0, 0, 0, 0, 0, 0, 0, 0, 0x36, 0, 0, 0, 0x28, 0, 0, 0,
                                0,    0,    0, 0, 0, 0, 0, 0, 0x01, 0, 0x20, 0, 0, 0, 0, 0, 0, 0,
                                0x10, 0,    0, 0, 0, 0, 0, 0, 0,    0, 0,    0, 0, 0, 0, 0, 0, 0};
    for (size_t b = 0; b < (size_t)w * h * 4; b += 4) std::swap(image[b], image[b + 2]);
    const long size = (long)w * h * 4 + 54, height = -h;
    for (int k = 0; k < 4; ++k) {
        header[2 + k] = (unsigned char)((size >> (8 * k)) & 0xff);
        header[18 + k] = (unsigned char)(((long)w >> (8 * k)) & 0xff);
        header[22 + k] = (unsigned char)((height >> (8 * k)) & 0xff);
    }
    FILE* fp = std::fopen((std::string(filename) + ".bmp").c_str(), "wb");
    if (!fp) return -1;
    std::fwrite(header, 1, 54, fp);
    std::fwrite(image, 1, (size_t)w * h * 4, fp);
    std::fclose(fp);
    return 0;
}

class Camera {
   public:
    int resolution;
    std::string name;

    static float3 tone(float3 c, int curve) {
        float3 o;
        mrp_tone_map_host(&c.x, 1, curve, &o.x);
        return o;
    }
    static float3 None(float3 c) { return c; }
    static float3 Gamma(float3 c) { return tone(c, 1); }
    static float3 ACES(float3 c) { return tone(c, 2); }

    // core/camera.cu:163-189: forward = -normalize(ori), right =
    // normalize(cross(Y, forward)), up = cross(forward, right) (unit: 90 deg).
    Camera(VolumeRender& volume, std::string name = "test camera", float3 position = {0.75f, 0.75f, 0.75f},
           int resolution = 512)
        : resolution(resolution), name(std::move(name)), volume_(&volume) {
        SetPosition(position);
    }
    void SetPosition(float3 position) {
        ori_ = position;
        forward_ = detail::normalize(-ori_);
        right_ = detail::normalize(detail::cross({0, 1, 0}, forward_));
        up_ = detail::cross(forward_, right_);
        if (position.x == 0 && position.z == 0) {
            right_ = {1, 0, 0};
            up_ = position.y > 0 ? float3{0, 0, 1} : float3{0, 0, -1};
        }
    }
    float3 GetPosition() const { return ori_; }
    // this build's extras: the basis SetPosition computed
    float3 GetUp() const { return up_; }
    float3 GetRight() const { return right_; }
    void SetVolume(VolumeRender& v) { volume_ = &v; }

    void Render(float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1, float g = 0,
                int sampleNum = 1, ToneMapFunction tone = ACES, VolumeRender::RenderType rt = VolumeRender::PT,
                float exp = 1.0f) {
        const std::string affix = rt == VolumeRender::PT ? "_PT" : (rt == VolumeRender::RPNN ? "_RPNN" : "MRPNN");
        RenderToFile("./" + name + affix, lightDir, lightColor, alpha, multiScatter, g, sampleNum, tone, rt, exp);
    }

    // core/camera.cu:226-252: render, tone(x * exp), saturate, x 255, SaveBMP(path).
    void RenderToFile(std::string path, float3 lightDir, float3 lightColor = {1, 1, 1}, float alpha = 1,
                      int multiScatter = 1, float g = 0, int sampleNum = 1, ToneMapFunction tone = ACES,
                      VolumeRender::RenderType rt = VolumeRender::PT, float exp = 1.0f) {
        const auto img = volume_->Render(int2{resolution, resolution}, ori_, up_, right_, lightDir, rt, g, alpha,
                                         lightColor, multiScatter, sampleNum);
        std::vector<unsigned char> data((size_t)resolution * resolution * 4);
        auto sat = [](float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); };
        for (size_t i = 0; i < img.size(); ++i) {
            const float3 r = tone(img[i] * exp);
            data[4 * i] = (unsigned char)(sat(r.x) * 255);
            data[4 * i + 1] = (unsigned char)(sat(r.y) * 255);
            data[4 * i + 2] = (unsigned char)(sat(r.z) * 255);
            data[4 * i + 3] = 255;
        }
        SaveBMP(data.data(), resolution, resolution, path.c_str());
    }

    // core/camera.cu:255-257
    void Render(float3* target, Histogram* histo_buffer, unsigned int* target2, int2 size, int frame, float3 lightDir,
                float3 lightColor = {1, 1, 1}, float alpha = 1, int multiScatter = 1, float g = 0, int tone = 2,
                VolumeRender::RenderType rt = VolumeRender::PT, bool denoise = false) {
        volume_->Render(target, histo_buffer, target2, size, ori_, up_, right_, lightDir, lightColor, alpha,
                        multiScatter, g, frame, rt, tone, denoise);
    }

   private:
    float3 ori_{}, forward_{}, right_{}, up_{};
    VolumeRender* volume_;
};

}  // namespace internship
