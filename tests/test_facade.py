"""The C++ facade (include/mrp_volume_render.hpp): it compiles against the C
ABI with the reference's class interface, and (GPU) programs written like
the reference's examples render the same image (and BMP) as the CPU oracle
and answer the point and ray queries as the oracle does."""
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
    bmp = np.fromfile(str(tmp_path / "cam") + ".bmp", dtype=np.uint8)
    assert bmp.size == 54 + size * size * 4
    # the camera basis Camera::SetPosition computed (core/camera.cu:173-189), read back
    ori, up, right = np.fromfile(str(out) + ".cam", dtype=np.float32).reshape(3, 3).astype(np.float64)
    fwd = -ori / np.linalg.norm(ori)
    r_ref = np.cross([0.0, 1.0, 0.0], fwd)
    r_ref /= np.linalg.norm(r_ref)
    assert np.abs(right - r_ref).max() < 1e-6 and np.abs(up - np.cross(fwd, r_ref)).max() < 1e-6
    p = RenderParams(size, size, Vec3.of(ori), Vec3.of(up), Vec3.of(right), Vec3.of((0.0, 0.957826, 0.287348)),
                     Vec3.of((1, 1, 1)), Vec3.of((0.1, 0.1, 0.1)), Vec3.of((0.99, 0.99, 0.99)), 0.857, 64.0, 1.0,
                     multi, 0, spp, 0, 0, 0)
    grid = sphere_grid(res)
    cpu, _, _ = oracle.render(grid, oracle.max_density(grid), p)
    diff = np.abs(img.astype(np.float64) - cpu)
    print("max abs diff %.3g, exact %.3f" % (diff.max(), np.mean(img == cpu)), flush=True)
    assert np.array_equal(img, cpu)
    # RenderToFile's BMP (core/camera.cu:10-64,226-252): the same image, ACES,
    # x 255, BGRA bytes top-down, as the oracle's tone and pack give them
    want = oracle.tone_pack(cpu.reshape(-1, 3), 1.0, 2).astype("<u4").view(np.uint8)
    assert np.array_equal(bmp[54:], want)
    assert bmp[0] == 0x42 and bmp[1] == 0x4d and int.from_bytes(bmp[22:26].tobytes(), "little", signed=True) == -size


def _sample(d, res, uv, clamp):
    """`Sample` / `SampleClamp` (core/volume.cu:675-722) in FP64."""
    p = uv * res - 0.5
    i = np.floor(p).astype(np.int64)
    w = p - i
    out = np.zeros(len(uv))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix, iy, iz = i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz
                if clamp:
                    ix, iy, iz = np.clip(ix, 0, res - 1), np.clip(iy, 0, res - 1), np.clip(iz, 0, res - 1)
                    ok = np.ones(len(uv), bool)
                else:
                    ok = (ix >= 0) & (iy >= 0) & (iz >= 0) & (ix < res) & (iy < res) & (iz < res)
                v = np.zeros(len(uv))
                v[ok] = np.maximum(d[ix[ok], iy[ok], iz[ok]], 0)
                out += (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                    (w[:, 2] if dz else 1 - w[:, 2]) * v
    return out


QBIN = os.path.join(ROOT, "tests", "cpp", "_build", "facade_queries")


@pytest.mark.gpu
def test_gpu_facade_queries_match_oracle(tmp_path):
    """`DensityAtPosition`, `DensityAtUV`, `TrAtPosition`, `GetHGLut`,
    `GetRadiances` (PT and MRPNN), `GetTrs`, `GetTr` and `GetSamples` through
    the facade (core/volume.hpp:102-131) against the oracle's mips, tables and
    estimators on the same inputs and streams."""
    import internship_amd as ia
    from internship_amd import RenderParams, Vec3
    from internship_amd._lib import SampleQuery
    pre = str(tmp_path / "q")
    wpath = os.path.join(ROOT, "internship_amd", "data", "mrpnn_weights.bin")
    print("case: facade_queries, 48^3 grid, 96 points and rays", flush=True)
    r = subprocess.run([QBIN, pre, wpath], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = lambda ext, shape: np.fromfile(pre + "." + ext, dtype=np.float32).reshape(shape)
    n, res = 96, 48
    grid = rd("grid", (res, res, res))
    pos, dirs = rd("pos", (n, 3)), rd("dir", (n, 3))
    light = np.array([0.0, 0.957826, 0.287348], np.float32)
    alpha, g, seed = 64.0, 0.857, 11
    mips, mc = oracle.mips(grid)
    trs, tc = oracle.tr_mips(mc, oracle.normalize3(light), alpha)
    lut = oracle.hg_lut(g, seed=seed)
    # host samplers
    dens = rd("dens", (n, 11)).astype(np.float64)
    uv = pos.astype(np.float64) + 0.5
    for m in range(9):
        np.testing.assert_allclose(dens[:, m], _sample(mips[m], 256 >> m, uv, False), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dens[:, 9], 0.5 * (dens[:, 1] + dens[:, 2]), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(dens[:, 10], dens[:, 2], rtol=1e-6, atol=1e-7)
    assert dens[:, 0].max() > 0.3
    tr = rd("tr", (n, 8)).astype(np.float64)
    ln = light.astype(np.float64)
    inside = np.all(np.abs(pos) <= 0.5, axis=1)
    q = uv.copy()
    for i in np.where(~inside)[0]:  # outside: the point where the light ray enters the box
        off, _ = oracle.ray_box(pos[i:i + 1], light[None, :])
        q[i] = uv[i] + ln * off[0] if off[0] >= 0 else np.nan
    for m in range(8):
        want = _sample(trs[m], 128 >> m, np.nan_to_num(q), True)
        want[np.isnan(q[:, 0])] = 1.0
        np.testing.assert_allclose(tr[:, m], want, rtol=1e-5, atol=1e-6)
    assert (~inside).sum() > 5 and tr.min() < 0.9
    hl = rd("lut", (n,)).astype(np.float64)
    cs, an = np.cos(0.07 * np.arange(n, dtype=np.float32)), 0.011 * np.arange(n, dtype=np.float32)
    u, v = cs * 0.5 + 0.5, an / (np.pi * 60 / 180)
    u0, v0 = np.clip((u * 512).astype(int), 0, 511), np.clip((v * 512).astype(int), 0, 511)
    u1, v1 = np.clip(u0 + 1, 0, 511), np.clip(v0 + 1, 0, 511)
    fu, fv = u * 512 - np.floor(u * 512), v * 512 - np.floor(v * 512)
    want = (lut[v0, u0] * (1 - fu) + lut[v0, u1] * fu) * (1 - fv) + (lut[v1, u0] * (1 - fu) + lut[v1, u1] * fu) * fv
    np.testing.assert_allclose(hl, want, rtol=1e-4, atol=1e-6)
    # GPU estimators
    md = oracle.max_density(grid)
    rori, rdir = rd("rori", (n, 3)), rd("rdir", (n, 3))
    V = Vec3.of
    p = RenderParams(1, 1, V((0, 0, 0)), V((0, 1, 0)), V((1, 0, 0)), V(light), V((2.0, 1.0, 0.5)), V((0.1, 0.2, 0.3)),
                     V((0.99, 0.95, 0.9)), g, alpha, 1.0, 16, 0, 4, seed, ia.PT, 0)
    want, _ = oracle.radiances(grid, md, rori, rdir, p)
    got = rd("rad", (n, 3))
    assert np.array_equal(got, want[:, :3]) and np.mean(got[:, 0] != np.float32(0.1)) > 0.3
    p.render_type, p.spp_count = ia.MRPNN, 1
    want, _ = oracle.radiances(grid, md, rori, rdir, p, mips_concat=mc, tr_concat=tc, lut=lut,
                               weights=ia.load_reference_weights())
    got = rd("radnn", (n, 3)).astype(np.float64)
    rel = np.abs(got - want[:, :3]) / np.maximum(np.abs(want[:, :3]), 1e-12)
    assert np.all((rel <= 1e-5) | (np.abs(got - want[:, :3]) <= 1e-7))
    t, _ = oracle.shadow_terms(grid, md, pos, dirs, light, alpha, g, 8, seed=seed)
    want = np.float32(np.pi) * t
    got = rd("trs", (n, 3))
    np.testing.assert_allclose(got, np.stack([2.0 * want, want, 0.5 * want], axis=1), rtol=1e-6, atol=1e-9)
    t1, _ = oracle.shadow_terms(grid, md, pos[3:4], dirs[3:4], light, alpha, g, 64, seed=seed)
    assert np.array_equal(rd("tr1", (3,)), np.full(3, t1[0], np.float32))
    m = 24
    qs = (SampleQuery * m)()
    for i in range(m):
        # the program's FP32 arithmetic: 0.7f + 0.008f * i
        gi = np.float32(np.float32(0.7) + np.float32(np.float32(0.008) * np.float32(i)))
        qs[i] = SampleQuery(V(pos[i]), V(dirs[i]), V(light), np.float32(32.0 + 4.0 * i), gi, np.float32(0.99))
    want, _ = oracle.get_samples(grid, md, qs, (2.0, 1.0, 0.5), (0.1, 0.2, 0.3), 16, 8, seed=seed)
    assert np.array_equal(rd("smp", (m, 3)), want) and want.max() > 0
