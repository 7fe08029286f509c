"""CPU pins of the oracle's training-point sampler (`GetDesiredCountSample` +
`MeanFreePathSample`, dataGen.cu:174-216).

- an empty volume gives no points;
- in a very dense homogeneous medium every free flight ends within a few
  thousandths of where it starts, so every first point lies on the origin
  sphere of radius 0.5 (up to the 1/128 jitter), and the origins are uniform
  on it (each coordinate of a uniform point on a sphere is uniform in
  [-r, r]: Archimedes);
- the same medium pins the rates the reference fixes: later vertices lie
  closer than 1/8 to the previous one, so each is kept with probability 0.1
  (1 + 3 x 0.1 points per ray); a kept point is jittered with probability 0.5
  by a uniform offset in the ball of radius 1/128, so the fraction of first
  points off the sphere by more than 1e-3 is 0.5 P(|z| > 0.128) for z the
  radial component of a uniform point in the unit ball, P = 1.5 (1 - a) -
  0.5 (1 - a^3); a later point's view direction is the first ray direction
  with probability 0.75;
- the rays head inward: the first view direction is the ray direction
  normalize(r + normalize(-origin)), whose dot product with the origin is
  negative on average.
"""
import math

import numpy as np

import oracle


def test_empty_volume_gives_no_points():
    grid = np.zeros((16, 16, 16), np.float32)
    pos, view, lk = oracle.training_points(grid, 1e-5, 500, 64.0, 0.857, seed=1)
    assert lk == 0 and np.all(pos[..., 3] == 0)


def _dense(n_rays, seed):
    grid = np.ones((16, 16, 16), np.float32)
    return oracle.training_points(grid, 1.0, n_rays, 20000.0, 0.857, seed=seed)


def _binom_ok(k, n, p, z=4.0):
    return abs(k - n * p) <= z * math.sqrt(n * p * (1 - p))


def test_dense_medium_first_points_on_the_origin_sphere():
    n = 20000
    pos, view, _ = _dense(n, 3)
    assert np.all(pos[:, 0, 3] == 1)  # every ray scatters at once
    p0 = pos[:, 0, :3].astype(np.float64)
    r = np.linalg.norm(p0, axis=1)
    assert np.all(np.abs(r - 0.5) < 1.0 / 128 + 2e-3)
    # origins uniform on the sphere: each coordinate uniform in [-0.5, 0.5] (KS)
    for k in range(3):
        x = np.sort(p0[:, k] / r * 0.5 + 0.5)
        ks = np.max(np.maximum(np.arange(1, n + 1) / n - x, x - np.arange(n) / n))
        assert ks < 2.0 / math.sqrt(n), (k, ks)
    # the first view direction is the ray direction: heading inward on average
    v0 = view[:, 0, :3].astype(np.float64)
    assert np.abs(np.linalg.norm(v0, axis=1) - 1).max() < 1e-5
    assert np.mean(np.sum(v0 * p0, axis=1)) < -0.2


def test_dense_medium_keep_jitter_and_view_rates():
    n = 40000
    pos, view, _ = _dense(n, 4)
    kept = pos[..., 3].sum()
    # vertices 1..3 always exist (the medium stops every ray) and move < 1/8: kept w.p. 0.1
    assert _binom_ok(kept - n, 3 * n, 0.1), kept / n
    p0 = pos[:, 0, :3].astype(np.float64)
    off = np.abs(np.linalg.norm(p0, axis=1) - 0.5) > 1e-3
    a = 0.128  # 1e-3 relative to the 1/128 radius
    p_off = 0.5 * (1.5 * (1 - a) - 0.5 * (1 - a ** 3))
    assert _binom_ok(off.sum(), n, p_off), (off.mean(), p_off)
    later = pos[:, 1:, 3] == 1
    same = np.all(view[:, 1:, :3] == view[:, :1, :3], axis=-1)
    assert _binom_ok(same[later].sum(), later.sum(), 0.75), same[later].mean()


def test_deterministic_and_thread_independent():
    from tests.conftest import sphere_grid
    grid = sphere_grid(32, radius=0.4)
    a = oracle.training_points(grid, 1.0, 3000, 64.0, 0.8, seed=9, threads=1)
    b = oracle.training_points(grid, 1.0, 3000, 64.0, 0.8, seed=9, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
