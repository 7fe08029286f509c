"""CPU pins of the procedural cloud generator (the synthetic inputs of the
benchmark and training configurations; the reference's own noise,
core/perlin_noise.cuh:4-7, is a stub whose .cu does not compile).

The GPU fill and the oracle fill share their source, so comparing them shows
only that both compile the same arithmetic.  These checks pin the generator
itself:
- Perlin's improved gradient noise is exactly zero at integer lattice
  points, for every octave whose frequency keeps the point on the lattice.
  With the base frequency at 2 * res every voxel centre (2i+1)/(2 res) is a
  lattice point of all octaves, so the cloud reduces to the closed form
  clamp((0.5 - t)/(1 - t), 0, 1) * clamp((0.5 - |p - c|)/0.2, 0, 1);
- densities lie in [0, 1] and vanish outside the falloff ball;
- the seed selects the permutation: the same seed gives the same cloud, and
  different seeds different ones;
- the 128^3 seed-1 cloud (the small version of the benchmark cloud) matches
  the committed checksum fixture (tests/golden/cloud128_seed1.json), which
  fixes the remap and falloff choices.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _closed_form(res, threshold):
    c = (2 * np.arange(res) + 1) / (2.0 * res)
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    fall = np.clip((0.5 - r) / 0.2, 0.0, 1.0)
    d = np.clip((0.5 - threshold) / (1.0 - threshold), 0.0, 1.0)
    return d * fall


@pytest.mark.parametrize("res,threshold", [(8, 0.2), (16, 0.4), (12, 0.0)])
def test_noise_vanishes_on_the_lattice(res, threshold):
    got = oracle.generate_cloud(res, seed=7, threshold=threshold, octaves=5, base_freq=2.0 * res)
    np.testing.assert_allclose(got, _closed_form(res, threshold), rtol=2e-6, atol=1e-7)
    assert got.max() > 0.1


def test_off_lattice_the_noise_varies():
    # the same grid at a frequency that puts no voxel centre on the lattice
    res = 16
    got = oracle.generate_cloud(res, seed=7, threshold=0.0, octaves=1, base_freq=3.0)
    base = _closed_form(res, 0.0)
    inside = base > 0.49  # falloff 1: density = (n + 1) / 2
    n = 2.0 * got[inside] - 1.0
    assert np.abs(n).max() <= 1.0
    assert n.std() > 0.05 and abs(n.mean()) < 0.2


def test_range_falloff_and_seeds():
    res = 32
    a = oracle.generate_cloud(res, seed=1)
    b = oracle.generate_cloud(res, seed=1)
    c = oracle.generate_cloud(res, seed=2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() >= 0.0 and a.max() <= 1.0
    outside = _closed_form(res, 0.0) == 0.0
    assert outside.any() and not a[outside].any()
    assert np.count_nonzero(a) > 0.1 * a.size


def test_cloud128_matches_checksum_fixture():
    with open(os.path.join(HERE, "golden", "cloud128_seed1.json")) as f:
        fix = json.load(f)
    g = oracle.generate_cloud(fix["res"], fix["seed"])
    assert hashlib.sha256(np.ascontiguousarray(g, np.float32).tobytes()).hexdigest() == fix["sha256"]
    assert int(np.count_nonzero(g)) == fix["nonzero"]
    for key, v in fix["samples"].items():
        i = tuple(int(t) for t in key.split(","))
        assert g[i] == np.float32(v)
