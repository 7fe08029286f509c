"""CPU pins of the oracle's point and ray queries (`GetRadiances`, `GetTr`,
`GetTrs`; core/volume.cu:302-402,630-673, core/render.cuh:200-258,317-341).

- a ray that misses the box returns the sky exactly with distance -1; an empty
  volume returns the sky;
- albedo 0: the mean radiance of a ray through a homogeneous box is the sky
  attenuated by Beer-Lambert;
- one bounce: the mean radiance is the FP64 quadrature of single scattering
  plus the attenuated sky, for the plain tracer and, with an all-zero
  network, for `NNPredict`, so both estimate the same thing;
- shadow terms in a homogeneous box: the mean of Tr * HG is
  exp(-sigma d_exit) * HG(dir . light) (ratio tracking is unbiased).
"""
import math

import numpy as np

import internship_amd as ia
import oracle
from tests.test_oracle_mrpnn_render import ALBEDO, LIGHT_COLOR, SKY, _single_scatter_quadrature

N_WEIGHTS = 49673


def _params(alpha, light, g=0.857, multi=1, spp=1, rt=ia.PT, seed=3):
    V = ia.Vec3.of
    return ia.RenderParams(1, 1, V((0, 0, 0)), V((0, 1, 0)), V((1, 0, 0)), V(light), V(LIGHT_COLOR), V(SKY), V(ALBEDO),
                           g, alpha, 1.0, multi, 0, spp, seed, rt, 0)


def _central(n):
    return np.tile([[0.0, 0.0, 2.5]], (n, 1)), np.tile([[0.0, 0.0, -1.0]], (n, 1))


def test_missing_rays_and_empty_volume_are_sky():
    grid = np.ones((8, 8, 8), np.float32)
    out, lk = oracle.radiances(grid, 1.0, [[0, 0, 2.5]], [[0, 1, 0]], _params(5.0, (0, 1, 0), multi=8, spp=4))
    assert lk == 0 and np.array_equal(out[0], np.float32(list(SKY) + [-1.0]))
    empty = np.zeros((8, 8, 8), np.float32)
    o, d = _central(16)
    out, _ = oracle.radiances(empty, 1e-5, o, d, _params(5.0, (0, 1, 0), multi=8, spp=2))
    np.testing.assert_allclose(out[:, :3], np.tile(SKY, (16, 1)), rtol=1e-6)
    assert np.all(out[:, 3] == -1)


def test_albedo_zero_is_beer_lambert():
    grid = np.ones((16, 16, 16), np.float32)
    alpha = 2.0
    p = _params(alpha, (0, 1, 0), multi=4, spp=64)
    p.scatter_rate = ia.Vec3.of((0, 0, 0))
    o, d = _central(2000)
    out, _ = oracle.radiances(grid, 1.0, o, d, p)
    tau = alpha * (0.4999 + 0.5)
    T = math.exp(-tau)
    for c in range(3):
        sigma = SKY[c] * math.sqrt(T * (1 - T) / (2000 * 64))
        assert abs(out[:, c].mean() - SKY[c] * T) < 5 * sigma


def test_one_bounce_pt_and_zero_network_match_single_scattering():
    alpha, g = 1.5, 0.857
    light = (0.0, 0.6, -0.8)
    grid = np.ones((16, 16, 16), np.float32)
    escape, single, hg = _single_scatter_quadrature(grid, alpha, g, light)
    n = 40000
    o, d = _central(n)
    pt, _ = oracle.radiances(grid, 1.0, o, d, _params(alpha, light, g, multi=1, spp=4))
    _, mc = oracle.mips(grid)
    _, tc = oracle.tr_mips(mc, light, alpha)
    nn, _ = oracle.radiances(grid, 1.0, o, d, _params(alpha, light, g, rt=ia.MRPNN), mips_concat=mc, tr_concat=tc,
                             lut=oracle.hg_lut(g, seed=0), weights=np.zeros(N_WEIGHTS, np.float32))
    for c in range(3):
        expected = SKY[c] * escape + LIGHT_COLOR[c] * ALBEDO[c] * single
        bound = max(SKY[c], LIGHT_COLOR[c] * ALBEDO[c] * hg)
        for name, got, paths in (("pt", pt[:, c].mean(), 4 * n), ("mrpnn", nn[:, c].mean(), n)):
            sigma = bound * 0.5 / math.sqrt(paths)
            assert abs(got - expected) < 5 * sigma + 2e-3 * expected, (name, c, got, expected)
    # first-vertex distances: exponential with rate alpha, truncated at the far face
    t = nn[nn[:, 3] >= 0, 3] - (2.5 - 0.4999)
    assert abs(t.mean() - (1 / alpha - (1.0 / (math.exp(alpha) - 1)))) < 0.02


def test_shadow_terms_homogeneous_box():
    grid = np.ones((16, 16, 16), np.float32)
    alpha, g = 1.2, 0.6
    light = np.array([0.0, 0.6, 0.8])
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.4, 0.4, (6, 3))
    dirs = rng.normal(size=(6, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    spp = 20000
    got, lk = oracle.shadow_terms(grid, 1.0, pos, dirs, light, alpha, g, spp, seed=4)
    assert lk > 0
    for i in range(6):
        d_exit = min(((0.5 if light[k] > 0 else -0.5) - pos[i, k]) / light[k] for k in range(3) if light[k] != 0)
        mu = float(dirs[i] @ light)
        hg = (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5 * 0.25
        T = math.exp(-alpha * d_exit)
        # per estimate values lie in [0, hg]
        assert abs(got[i] - T * hg) < 5 * hg * 0.5 / math.sqrt(spp) + 1e-3 * hg, (i, got[i], T * hg)
