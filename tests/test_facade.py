"""The C++ facade (include/mrp_volume_render.hpp): it compiles against the C
ABI with the reference's class interface, and (GPU) a program written like
the reference's examples renders the same image as the CPU oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from tests.conftest import sphere_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "facade_render")


def test_facade_header_compiles_with_reference_style_code(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "mrp_volume_render.hpp"
using namespace internship;
int main() {
    VolumeRender volume(64);
    volume.SetDatas([](int x, int y, int z, float u, float v, float w) { return u * v * w; });
    volume.Update();
    volume.Update_TR(make_float3(0, 1, 0), 64.0f);
    volume.UpdateHGLut(0.857f);
    volume.SetScatterRate(0.99f);
    volume.SetExposure(1.0f);
    volume.SetCheckboard(false);
    Camera cam(volume, "test camera", make_float3(0.75f, 0.75f, 0.75f), 64);
    cam.SetPosition(make_float3(0, 1, 0));
    auto img = volume.Render(int2{8, 8}, make_float3(0, 0, 2), make_float3(0, 1, 0), make_float3(1, 0, 0),
                             make_float3(0, 1, 0), VolumeRender::MRPNN, 0.857f, 64.0f);
    float3* target = nullptr; Histogram* histo = nullptr; unsigned* bgra = nullptr;
    cam.Render(target, histo, bgra, int2{8, 8}, 0, make_float3(0, 1, 0));
    return img.size() == 64 ? 0 : 1;
}
''')
    gxx = shutil.which("g++") or "g++"
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_facade_test_program_is_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_gpu_facade_render_matches_oracle(tmp_path):
    from internship_amd import RenderParams, Vec3
    size, spp, multi, res = 24, 2, 16, 48
    out = tmp_path / "img.raw"
    print("case: facade_render %d^3 sphere, %dx%d, %d spp" % (res, size, size, spp), flush=True)
    r = subprocess.run([BIN, str(out), str(res), str(size), str(spp), str(multi), str(tmp_path / "cam")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    img = np.fromfile(out, dtype=np.float32).reshape(size, size, 3)
    assert os.path.getsize(str(tmp_path / "cam") + ".bmp") == 54 + size * size * 4
    ori, up, right = np.fromfile(str(out) + ".cam", dtype=np.float32).reshape(3, 3)
    fwd = -ori / np.linalg.norm(ori)
    assert abs(np.linalg.norm(right) - 1) < 1e-6 and abs(right @ fwd) < 1e-6 and abs(up @ right) < 1e-6
    p = RenderParams(size, size, Vec3.of(ori), Vec3.of(up), Vec3.of(right), Vec3.of((0.0, 0.957826, 0.287348)),
                     Vec3.of((1, 1, 1)), Vec3.of((0.1, 0.1, 0.1)), Vec3.of((0.99, 0.99, 0.99)), 0.857, 64.0, 1.0,
                     multi, 0, spp, 0, 0, 0)
    grid = sphere_grid(res)
    cpu, _, _ = oracle.render(grid, oracle.max_density(grid), p)
    diff = np.abs(img.astype(np.float64) - cpu)
    print("max abs diff %.3g, exact %.3f" % (diff.max(), np.mean(img == cpu)), flush=True)
    assert np.array_equal(img, cpu)
