// Renders a small scene through the C++ façade (include/mrp_volume_render.hpp)
// the way reference code does (README.md:74-92, main.cu:9-47) and writes the
// FP32 image for the parity test to compare with the oracle.
//   facade_render <out.raw> <resolution> <image size> <spp> <multi scatter> [bmp path]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "mrp_volume_render.hpp"

using namespace internship;

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s out.raw res size spp multi [bmp]\n", argv[0]);
        return 2;
    }
    const int res = std::atoi(argv[2]), size = std::atoi(argv[3]), spp = std::atoi(argv[4]),
              multi = std::atoi(argv[5]);
    VolumeRender volume(res);
    // the README's sphere: density 1 inside radius 0.25 of the centre
    volume.SetDatas([](int, int, int, float u, float v, float w) {
        const float dx = u - 0.5f, dy = v - 0.5f, dz = w - 0.5f;
        return std::sqrt(dx * dx + dy * dy + dz * dz) < 0.25f ? 1.0f : 0.0f;
    });
    volume.Update();
    volume.SetScatterRate(0.99f);
    volume.SetSky(make_float3(0.1f, 0.1f, 0.1f));
    Camera camera(volume, "facade", make_float3(1.5f, 1.2f, 1.4f), size);
    const float3 light = make_float3(0.0f, 0.957826f, 0.287348f);
    if (argc > 6)
        camera.RenderToFile(argv[6], light, make_float3(1, 1, 1), 64.0f, multi, 0.857f, spp, Camera::ACES,
                            VolumeRender::PT);
    // the same camera's rays (the basis Camera::SetPosition computed), the raw image
    const float3 ori = camera.GetPosition(), up = camera.GetUp(), right = camera.GetRight();
    const auto img = volume.Render(int2{size, size}, ori, up, right, light, VolumeRender::PT, 0.857f, 64.0f,
                                   make_float3(1, 1, 1), multi, spp);
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    std::fwrite(img.data(), sizeof(float3), img.size(), f);
    std::fclose(f);
    // the camera vectors used, for the comparison
    const float3 cam[3] = {ori, up, right};
    f = std::fopen((std::string(argv[1]) + ".cam").c_str(), "wb");
    if (!f) return 1;
    std::fwrite(cam, sizeof(float3), 3, f);
    std::fclose(f);
    std::printf("max density %g, %zu pixels\n", volume.max_density, img.size());
    return 0;
}
