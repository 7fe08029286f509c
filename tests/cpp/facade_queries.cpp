// Point and ray queries through the C++ façade (include/mrp_volume_render.hpp)
// the way reference code calls them (dataGen.cu:273-302, core/volume.hpp:
// 102-131).  Writes its grid, inputs and results as raw FP32 for the parity
// test to recompute with the oracle.
//   facade_queries <out prefix> <weights.bin>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mrp_volume_render.hpp"

using namespace internship;

static bool dump(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s out_prefix weights.bin\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int res = 48;
    auto fill = [](int, int, int, float u, float v, float w) {
        const float dx = u - 0.5f, dy = v - 0.5f, dz = w - 0.52f;
        const float r = std::sqrt(dx * dx + dy * dy + dz * dz);
        return r < 0.4f ? (1.0f - r / 0.4f) * (0.5f + 0.5f * u) : 0.0f;
    };
    std::vector<float> grid((size_t)res * res * res);
    for (int i = 0; i < res; ++i)
        for (int j = 0; j < res; ++j)
            for (int k = 0; k < res; ++k)
                grid[((size_t)i * res + j) * res + k] =
                    fill(i, j, k, (float)((i + 0.5) / res), (float)((j + 0.5) / res), (float)((k + 0.5) / res));
    VolumeRender volume(res);
    volume.SetDatas(fill);
    volume.Update();
    volume.SetScatterRate(make_float3(0.99f, 0.95f, 0.9f));
    volume.SetSky(make_float3(0.1f, 0.2f, 0.3f));
    volume.SetSeed(11);
    volume.SetMlpPrecision(0);
    if (!volume.LoadNetwork(argv[2])) return 1;
    const float3 light = make_float3(0.0f, 0.957826f, 0.287348f), lc = make_float3(2.0f, 1.0f, 0.5f);
    const float alpha = 64.0f, g = 0.857f;
    volume.Update_TR(light, alpha);
    volume.UpdateHGLut(g);

    // 64 points on a spiral inside the cloud, 32 outside the box; directions about them
    const int n = 96;
    std::vector<float3> pos(n), dir(n);
    for (int i = 0; i < n; ++i) {
        const float t = (i + 0.5f) / n, a = 2.399963f * i;
        if (i < 64) {
            const float r = 0.33f * (i + 0.5f) / 64.0f;
            pos[i] = make_float3(r * std::cos(a), 0.25f * std::sin(3.0f * a) * r / 0.33f, 0.02f + r * std::sin(a));
        } else {
            pos[i] = make_float3(0.7f * std::cos(a), -0.3f + 0.9f * (t - 0.66f), 0.7f * std::sin(a));
        }
        const float b = 1.7f * i;
        dir[i] = make_float3(std::cos(b) * std::sin(0.3f + t), std::cos(0.3f + t), std::sin(b) * std::sin(0.3f + t));
    }
    std::vector<float> dens, tr, lut;
    for (int i = 0; i < n; ++i) {
        for (int m = 0; m < 9; ++m) dens.push_back(volume.DensityAtPosition(m, pos[i]));
        dens.push_back(volume.DensityAtPosition(1.5f, pos[i]));
        dens.push_back(volume.DensityAtUV(2, pos[i] + 0.5f));
        for (int m = 0; m < 8; ++m) tr.push_back(volume.TrAtPosition(m, pos[i], light));
        lut.push_back(volume.GetHGLut(std::cos(0.07f * i), 0.011f * i));
    }
    // rays from outside towards the box
    std::vector<float3> rori(n), rdir(n);
    for (int i = 0; i < n; ++i) {
        rori[i] = make_float3(2.0f, 0.3f * std::sin(0.5f * i), 0.3f * std::cos(0.7f * i));
        rdir[i] = make_float3(-1.0f, 0.2f * std::sin(0.9f * i), 0.2f * std::cos(1.3f * i));
    }
    const auto rad = volume.GetRadiances(rori, rdir, light, lc, alpha, 16, g, 4, VolumeRender::PT);
    const auto rad_nn = volume.GetRadiances(rori, rdir, light, lc, alpha, 16, g, 1, VolumeRender::MRPNN);
    const auto trs = volume.GetTrs(alpha, pos, dir, light, lc, g, 8);
    const float3 tr1 = volume.GetTr(pos[3], dir[3], light, alpha, g, 40);
    const int m = 24;
    std::vector<float> q_alpha(m), q_g(m), q_sc(m);
    std::vector<float3> q_pos(pos.begin(), pos.begin() + m), q_dir(dir.begin(), dir.begin() + m), q_l(m, light);
    for (int i = 0; i < m; ++i) {
        q_alpha[i] = 32.0f + 4.0f * i;
        q_g[i] = 0.7f + 0.008f * i;
        q_sc[i] = 0.99f;
    }
    const auto smp = volume.GetSamples(q_alpha, q_pos, q_dir, q_l, q_g, q_sc, lc, 16, 8);
    const bool ok = dump(out + ".grid", grid.data(), grid.size() * 4) && dump(out + ".pos", pos.data(), n * 12) &&
                    dump(out + ".dir", dir.data(), n * 12) && dump(out + ".dens", dens.data(), dens.size() * 4) &&
                    dump(out + ".tr", tr.data(), tr.size() * 4) && dump(out + ".lut", lut.data(), lut.size() * 4) &&
                    dump(out + ".rori", rori.data(), n * 12) && dump(out + ".rdir", rdir.data(), n * 12) &&
                    dump(out + ".rad", rad.data(), n * 12) && dump(out + ".radnn", rad_nn.data(), n * 12) &&
                    dump(out + ".trs", trs.data(), n * 12) && dump(out + ".tr1", &tr1, 12) &&
                    dump(out + ".smp", smp.data(), m * 12);
    std::printf("max density %g, %d points, %s\n", volume.max_density, n, ok ? "written" : "write failed");
    return ok ? 0 : 1;
}
