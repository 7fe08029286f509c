"""Pins of the CPU oracle's leaf functions (no GPU).

Every check here compares the oracle with something independent of it:
published known-answer vectors, libm in float64, closed forms, or brute force.
"""
import math

import numpy as np
import pytest

import oracle


# ---------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 philox4x32-10 known-answer vectors (Salmon et al., SC'11).
    assert [hex(v) for v in oracle.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(v) for v in oracle.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(v) for v in oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                          [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_uniforms_range_and_moments():
    u = np.concatenate([oracle.uniforms(i, 3, 0, 12345, 4096) for i in range(16)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.005
    assert abs(u.var() - 1.0 / 12.0) < 0.002
    # all draws are multiples of 2^-24
    k = u.astype(np.float64) * 2 ** 24
    assert np.all(k == np.round(k))


def test_uniform_stream_is_philox_words():
    # draw n = ((word n%4 of block n/4) >> 8) + 1) * 2^-24 with counter (id, sample, block, tag)
    u = oracle.uniforms(7, 9, 0, 5, 8)
    for n in range(8):
        w = oracle.philox([7, 9, n // 4, 0], [5, 0])[n % 4]
        assert u[n] == np.float32(((int(w) >> 8) + 1) * 2.0 ** -24)


# ---------------------------------------------------------------- scalar math vs libm
def _ulps(a, ref):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    spacing = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(a - ref) / np.maximum(spacing, np.finfo(np.float32).tiny)


def test_log_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 1, 200000), np.exp(rng.uniform(-80, 80, 200000)),
                        np.arange(1, 2 ** 24 + 1, 997) * 2.0 ** -24]).astype(np.float32)
    got = oracle.math("log", x)
    assert _ulps(got, np.log(x.astype(np.float64))).max() <= 2.0
    # special values
    sp = oracle.math("log", np.array([0.0, 1.0, np.inf, -1.0, 1e-45], dtype=np.float32))
    assert sp[0] == -np.inf and sp[1] == 0.0 and sp[2] == np.inf and np.isnan(sp[3])
    assert abs(sp[4] - math.log(1.401298464324817e-45)) < 1e-4


def test_exp_accuracy():
    rng = np.random.default_rng(1)
    x = rng.uniform(-87, 88, 400000).astype(np.float32)
    got = oracle.math("exp", x)
    assert _ulps(got, np.exp(x.astype(np.float64))).max() <= 2.0
    assert oracle.math("exp", np.array([0.0], dtype=np.float32))[0] == 1.0
    assert oracle.math("exp", np.array([100.0], dtype=np.float32))[0] == np.inf
    assert oracle.math("exp", np.array([-200.0], dtype=np.float32))[0] == 0.0


def test_sincos_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(0, 2 * math.pi, 200000), rng.uniform(-20, 20, 100000)]).astype(np.float32)
    s = oracle.math("sin", x).astype(np.float64)
    c = oracle.math("cos", x).astype(np.float64)
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(xd)).max() < 2.5e-7
    assert np.abs(c - np.cos(xd)).max() < 2.5e-7


def test_atan_acos_accuracy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-5, 5, 200000), np.tan(rng.uniform(-1.5, 1.5, 100000))]).astype(np.float32)
    assert _ulps(oracle.math("atan", x), np.arctan(x.astype(np.float64))).max() <= 3.0
    y = rng.uniform(-1, 1, 200000).astype(np.float32)
    assert np.abs(oracle.math("acos", y).astype(np.float64) - np.arccos(y.astype(np.float64))).max() < 5e-7
    assert oracle.math("acos", np.array([1.0], dtype=np.float32))[0] == 0.0
    assert abs(oracle.math("acos", np.array([-1.0], dtype=np.float32))[0] - math.pi) < 3e-7


# ---------------------------------------------------------------- Henyey-Greenstein
@pytest.mark.parametrize("g", [0.0, 0.3, 0.857, -0.5])
def test_hg_phase_closed_form_and_normalisation(g):
    mu = np.linspace(-1, 1, 20001).astype(np.float32)
    got = oracle.math("hg", mu, np.full_like(mu, g)).astype(np.float64)
    gf = float(np.float32(g))
    ref = (1 - gf * gf) / (1 + gf * gf - 2 * gf * mu.astype(np.float64)) ** 1.5 * 0.25
    # float32 cancellation in 1+g^2-2g*mu near the forward peak bounds the agreement
    assert np.abs(got / ref - 1).max() < 2e-5
    # no 1/pi factor (core/vector.cu:208): integral over the sphere is pi
    integral = 2 * math.pi * np.trapezoid(got, mu.astype(np.float64))
    assert abs(integral - math.pi) < 2e-3 * math.pi


@pytest.mark.parametrize("g", [0.0, 0.5, 0.857, -0.7])
def test_hg_inverse_cdf(g):
    s = np.linspace(1e-4, 1, 5001).astype(np.float32)
    mu = oracle.math("hg_cos", s, np.full_like(s, g)).astype(np.float64)
    assert np.all(mu >= -1 - 1e-5) and np.all(mu <= 1 + 1e-5)
    if abs(g) < 1e-4:
        cdf = (mu + 1) / 2
    else:
        cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
    assert np.abs(cdf - s).max() < 2e-4


def test_hg_sample_dir_distribution():
    rng = np.random.default_rng(4)
    n = 20000
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    e0 = rng.uniform(1e-6, 1, n).astype(np.float32)
    e1 = rng.uniform(1e-6, 1, n).astype(np.float32)
    g = 0.857
    d = oracle.hg_sample_dir(e0, e1, v.astype(np.float32), g).astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    cos_expected = oracle.math("hg_cos", e0, np.full_like(e0, g)).astype(np.float64)
    assert np.abs(np.sum(d * v, axis=1) - cos_expected).max() < 2e-5


# ---------------------------------------------------------------- ray-box
def _slab_np(p, d, half):
    inv = np.where(d == 0, 1e9, 1.0 / np.where(d == 0, 1.0, d))
    t0 = (-half - p) * inv
    t1 = (half - p) * inv
    tmin = np.minimum(t0, t1).max(axis=1)
    tmax = np.maximum(t0, t1).min(axis=1)
    return tmin, tmax


def test_ray_box_against_slab_method():
    rng = np.random.default_rng(5)
    n = 20000
    p = rng.uniform(-2, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    off, dist = oracle.ray_box(p.astype(np.float32), d.astype(np.float32))
    tmin, tmax = _slab_np(p, d, 0.4999)
    hit = tmax >= tmin
    ref_off = np.where(hit, np.maximum(tmin, 0), -1)
    ok = np.abs(off - ref_off) < 1e-4
    assert ok.mean() > 0.999
    _, tmax5 = _slab_np(p, d, 0.5)
    # FP32 slab arithmetic: error scales with |p|/|d_i| of the limiting axis
    scale = 1.0 + np.abs(p).max(axis=1) / np.maximum(np.abs(d).min(axis=1), 1e-9)
    assert np.all(np.abs(dist - tmax5) <= 1e-6 * scale + 1e-6)
    # closed forms
    o, dd = oracle.ray_box(np.array([[2, 0, 0], [0, 0, 0]], np.float32), np.array([[-1, 0, 0], [0, 1, 0]], np.float32))
    assert abs(o[0] - 1.5001) < 1e-6 and o[1] == 0.0 and dd[1] == 0.5


# ---------------------------------------------------------------- trilinear density
def _trilinear_np(grid, pos):
    res = grid.shape[0]
    t = (pos.astype(np.float64) + 0.5) * res - 0.5
    i0 = np.floor(t).astype(np.int64)
    w = t - i0
    out = np.zeros(pos.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix, iy, iz = i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz
                inside = (ix >= 0) & (iy >= 0) & (iz >= 0) & (ix < res) & (iy < res) & (iz < res)
                val = np.where(inside, grid[np.clip(ix, 0, res - 1), np.clip(iy, 0, res - 1), np.clip(iz, 0, res - 1)], 0)
                wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * (w[:, 2] if dz else 1 - w[:, 2])
                out += wt * val
    return out


def test_density_trilinear_border():
    rng = np.random.default_rng(6)
    grid = rng.uniform(0, 1, (8, 8, 8)).astype(np.float32)
    pos = rng.uniform(-0.7, 0.7, (50000, 3)).astype(np.float32)
    got = oracle.density(grid, pos)
    ref = _trilinear_np(grid, pos)
    assert np.abs(got - ref).max() < 3e-6
    # voxel centres reproduce the grid exactly
    c = (np.arange(8) + 0.5) / 8 - 0.5
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    centres = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    assert np.array_equal(oracle.density(grid, centres), grid.reshape(-1))


def test_max_density_floor():
    assert oracle.max_density(np.zeros((4, 4, 4), np.float32)) == np.float32(1e-5)
    g = np.zeros((4, 4, 4), np.float32)
    g[1, 2, 3] = 0.75
    assert oracle.max_density(g) == np.float32(0.75)
