"""The environment map (`SetHDRI` / `SkyBox`, core/volume.cu:1426-1458,
core/render.cuh:57-59), the light choice of MRPNN renders with a map
(core/volume.cu:417-432) and the Radiance RGBE reader, on the CPU.

- a constant map c with exposure e is the constant sky c e, texel for texel
  and through whole oracle renders (bit for bit, plain and MRPNN);
- the lookup is a bilinear, wrap-addressed read at u = atan2(-z, x)/(2 pi) +
  1/2, v = acos(y)/pi on texel centres: an FP64 numpy lookup of a smooth map
  agrees, the seam at u = 0 / 1 blends the first and last columns, and the
  axis directions land where the formulas put them;
- the light choice: with probability 4e/(4e + max L) the light is the
  uniform-sphere image of the R2 point n (frac(n/g + 1/2), frac(n/g^2 + 1/2),
  g the plastic number, all in FP32 as the reference's float g gives them)
  with the map's colour times 4e/rate, else L/(1-rate);
- RGBE: flat and run-length-encoded files written here decode to
  byte * 2^(e - 136) (G. Ward's definition), e = 0 black.
"""
import math

import numpy as np
import pytest

import internship_amd as ia
import oracle
from tests.conftest import config1_params, sphere_grid


@pytest.fixture(autouse=True)
def _clear_map():
    yield
    oracle.set_hdri(None)


def _smooth_map(w=64, h=32):
    v = (np.arange(h) + 0.5)[:, None] / h
    u = (np.arange(w) + 0.5)[None, :] / w
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = 1.0 + 0.5 * np.cos(2 * np.pi * u) * np.sin(np.pi * v)
    img[..., 1] = 0.5 + 0.3 * v + 0 * u
    img[..., 2] = 0.2 + 0.1 * np.sin(2 * np.pi * u)
    img[..., 3] = 1.0
    return img


def _lookup64(img, d, exposure):
    d = np.asarray(d, np.float64)
    h, w = img.shape[:2]
    u = np.arctan2(-d[:, 2], d[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(d[:, 1], -1, 1)) / np.pi
    tx, ty = u * w - 0.5, v * h - 0.5
    fx, fy = np.floor(tx), np.floor(ty)
    wx, wy = (tx - fx)[:, None], (ty - fy)[:, None]
    x0, x1 = fx.astype(int) % w, (fx.astype(int) + 1) % w
    y0, y1 = fy.astype(int) % h, (fy.astype(int) + 1) % h
    p = img[..., :3].astype(np.float64)
    top = p[y0, x0] * (1 - wx) + p[y0, x1] * wx
    bot = p[y1, x0] * (1 - wx) + p[y1, x1] * wx
    return (top * (1 - wy) + bot * wy) * exposure


def test_constant_map_is_the_constant_sky():
    c = np.array([0.3, 0.6, 0.9], np.float32)
    img = np.zeros((16, 32, 4), np.float32)
    img[..., :3] = c
    oracle.set_hdri(img, 0.5)
    rng = np.random.default_rng(1)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.array_equal(oracle.env(d), np.tile(c * np.float32(0.5), (500, 1)))
    # whole renders: the map c with exposure 0.5 == the constant sky c * 0.5
    grid = sphere_grid(32)
    p = config1_params(size=12, spp=3, multi=16, seed=4)
    with_map, lk1, _ = oracle.render(grid, 1.0, p)
    oracle.set_hdri(None)
    p.sky = ia.Vec3.of(c * np.float32(0.5))
    with_sky, lk2, _ = oracle.render(grid, 1.0, p)
    assert lk1 == lk2 and np.array_equal(with_map, with_sky)


def test_lookup_matches_fp64_bilinear_wrap():
    img = _smooth_map()
    oracle.set_hdri(img, 2.0)
    rng = np.random.default_rng(2)
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    np.testing.assert_allclose(oracle.env(d), _lookup64(img, d, 2.0), rtol=2e-5, atol=2e-6)


def test_seam_and_axis_directions():
    w, h = 8, 4
    img = np.zeros((h, w, 4), np.float32)
    img[:, 0, 0] = 1.0
    img[:, w - 1, 0] = 3.0
    img[:, w // 2 - 1, 1] = 5.0
    img[:, w // 2, 1] = 7.0
    oracle.set_hdri(img, 1.0)
    # -x: atan2(-0, -1) = pi, u = 1 (and atan2(+0, -1)): between the last and the first column
    r = oracle.env([[-1.0, 0.0, 0.0], [-1.0, 0.0, -0.0]])
    assert np.allclose(r[:, 0], 2.0, atol=1e-5)
    # +x: u = 1/2, halfway between columns w/2 - 1 and w/2; v = 1/2
    r = oracle.env([[1.0, 0.0, 0.0]])
    assert abs(r[0, 1] - 6.0) < 1e-5 and abs(r[0, 0]) < 1e-6
    # -z: atan2(1, 0) = pi/2, u = 3/4 (column 5.5 -> 5 and 6: zero here)
    assert np.all(np.abs(oracle.env([[0.0, 0.0, -1.0]])) < 1e-6)


def test_env_light_choice():
    img = np.zeros((8, 16, 4), np.float32)
    img[..., :3] = [0.5, 1.0, 2.0]
    e = 0.75
    oracle.set_hdri(img, e)
    L, C = np.array([0.0, 1.0, 0.0]), np.array([3.0, 2.0, 1.0])
    rate = 4 * e / (4 * e + 3.0)
    g = 1.32471795724474602596
    f = np.float32
    for n in (0, 1, 7, 100):
        d, c = oracle.env_light(n, rate - 1e-3, L, C)
        # the R2 point in FP32, as the reference's float3 arithmetic forms it
        gf = f(g)
        px, py = f(f(1) / gf) * f(n) + f(0.5), f(f(1) / (gf * gf)) * f(n) + f(0.5)
        rx, ry = float(px - np.floor(px)), float(py - np.floor(py))
        phi, ct = 2 * np.pi * rx, 1 - 2 * ry
        want = np.array([np.sqrt(1 - ct * ct) * np.cos(phi), np.sqrt(1 - ct * ct) * np.sin(phi), ct])
        assert np.abs(d - want).max() < 2e-5, (n, d, want)
        np.testing.assert_allclose(c, np.array([0.5, 1.0, 2.0]) * 4 * e / rate, rtol=1e-5)
        d, c = oracle.env_light(n, rate + 1e-3, L, C)
        assert np.array_equal(d, L.astype(np.float32))
        np.testing.assert_allclose(c, C / (1 - rate), rtol=1e-5)


# ------------------------------------------------------------------- RGBE

def _rgbe_bytes(img):
    rgb = img[..., :3].astype(np.float64)
    m = rgb.max(axis=-1)
    mant, ex = np.frexp(m)
    scale = np.where(m > 1e-32, mant * 256.0 / np.where(m > 0, m, 1), 0)
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255)
    out[..., 3] = np.where(m > 1e-32, ex + 128, 0)
    return out


def _rle_channel(b):
    out = bytearray()
    i = 0
    while i < len(b):
        j = i
        while j < len(b) and j - i < 127 and b[j] == b[i]:
            j += 1
        if j - i >= 4:
            out += bytes([128 + (j - i), b[i]])
            i = j
        else:
            k = i
            while k < len(b) and k - i < 128 and not (k + 3 < len(b) and b[k] == b[k + 1] == b[k + 2] == b[k + 3]):
                k += 1
            k = max(k, i + 1)
            out += bytes([k - i]) + bytes(b[i:k])
            i = k
    return bytes(out)


def _write_hdr(path, px, rle):
    h, w = px.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y %d +X %d\n" % (h, w))
        for y in range(h):
            if rle:
                f.write(bytes([2, 2, w >> 8, w & 255]))
                for c in range(4):
                    f.write(_rle_channel(px[y, :, c].tobytes()))
            else:
                f.write(px[y].tobytes())


@pytest.mark.parametrize("rle", [False, True])
def test_rgbe_reader(tmp_path, rle):
    rng = np.random.default_rng(5)
    img = np.zeros((6, 40, 4), np.float32)
    img[..., :3] = rng.uniform(0, 30, (6, 40, 3)) * (rng.random((6, 40, 1)) < 0.8)
    img[2, 5:20, :3] = 4.0  # runs
    px = _rgbe_bytes(img)
    p = str(tmp_path / ("m_rle.hdr" if rle else "m.hdr"))
    _write_hdr(p, px, rle)
    got = ia.read_hdr(p)
    want = np.zeros_like(got)
    s = np.where(px[..., 3] > 0, np.ldexp(1.0, px[..., 3].astype(int) - 136), 0.0)
    want[..., :3] = px[..., :3] * s[..., None]
    want[..., 3] = 1.0
    assert got.shape == (6, 40, 4) and np.array_equal(got, want.astype(np.float32))
    assert np.abs(got[..., :3] - img[..., :3]).max() <= img.max() / 128  # 8-bit mantissas
