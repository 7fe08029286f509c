"""CPU pins of the oracle's serving composition (oracle.serve_view, the
restatement mrp_serve_views is checked against):

- with a threshold no pixel reaches, the estimate is the MRPNN offline render
  of samples 0 .. spp-1 bit for bit (the histogram's running mean is the
  renderer's running mean);
- the flag is |E[l^2] - E[l]^2| > threshold for l = sum over channels of
  r/(r+1), E the running mean of the frame path's histogram: recomputed in
  numpy float32 from the per-sample MRPNN values (each read back exactly from
  a one-sample render's FP64 sum);
- a flagged pixel is re-estimated from scratch with delta-tracking samples
  spp .. spp+K-1 on the offline tracer's streams: with K = 1 it equals the
  FP64 sum of oracle.render(spp_begin=spp, spp_count=1) exactly, with K = 3
  the FP64 mean of spp_begin=spp, spp_count=3 to FP32 rounding; its histogram
  holds K samples, an unflagged one spp;
- a larger threshold flags a subset of the pixels a smaller one flags.
"""
from fractions import Fraction

import numpy as np
import pytest

import internship_amd as ia
import oracle
from tests.conftest import config1_params

W, H = 24, 16
SPP = 3


@pytest.fixture(scope="module")
def scene():
    grid = oracle.generate_cloud(32, seed=5)
    md = oracle.max_density(grid)
    p = config1_params(size=(W, H), spp=SPP, multi=8, seed=7)
    p.render_type, p.mlp_precision = 2, 0
    p.light_color = ia.Vec3.of((1.5, 1.0, 0.8))
    _, mc = oracle.mips(grid)
    _, tc = oracle.tr_mips(mc, [p.light_dir.x, p.light_dir.y, p.light_dir.z], p.alpha)
    lut = oracle.hg_lut(p.g, seed=p.seed)
    w = ia.load_reference_weights()
    return grid, md, mc, tc, lut, w, p


def _serve(scene, threshold, k):
    grid, md, mc, tc, lut, w, p = scene
    return oracle.serve_view(grid, md, mc, tc, lut, w, p, SPP, threshold, k, threads=8)


def _mrpnn_samples(scene):
    """Per-sample MRPNN radiance (h, w, spp, 3), clamped at 0, exactly."""
    grid, md, mc, tc, lut, w, p = scene
    out = np.zeros((H, W, SPP, 3), np.float32)
    for s in range(SPP):
        q = config1_params(size=(W, H), spp=1, multi=8, seed=7, spp_begin=s)
        q.render_type, q.mlp_precision, q.light_color = 2, 0, p.light_color
        sums = np.zeros((H, W, 3), np.float64)
        oracle.render_mrpnn(grid, md, mc, tc, lut, w, q, target=np.zeros((H, W, 3), np.float32), sum64=sums, threads=8)
        out[:, :, s] = sums.astype(np.float32)
        assert np.array_equal(out[:, :, s].astype(np.float64), sums)
    return out


def test_no_fallback_is_the_mrpnn_render(scene):
    grid, md, mc, tc, lut, w, p = scene
    rgb, histo, flags, lk = _serve(scene, 1e30, 4)
    ref = oracle.render_mrpnn(grid, md, mc, tc, lut, w, p, threads=8)
    assert not flags.any() and lk == 0
    assert np.array_equal(rgb, ref)
    assert np.all(histo[..., 30] == SPP)  # sample count


def _fma32(a, b, c):
    """FP32 fma(a, b, c), rounded once (ties to even) from the exact value."""
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    f = np.float32(float(exact))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    err = [abs(Fraction(float(q)) - exact) for q in cands]
    best = min(err)
    picks = [q for q, e in zip(cands, err) if e == best]
    return picks[0] if len(picks) == 1 else [q for q in picks if (q.view(np.uint32) & 1) == 0][0]


def _lerp32(a, b, v):
    """The shared math's lerp: fma(v, b - a, a) in FP32."""
    out = np.empty_like(a)
    d = (b - a).astype(np.float32)
    for i in np.ndindex(a.shape):
        out[i] = _fma32(v, d[i], a[i])
    return out


def test_flag_is_the_histogram_variance(scene):
    r = _mrpnn_samples(scene)
    f = np.float32
    t = r / (r + f(1))
    lum = (t[..., 0] + t[..., 1]) + t[..., 2]
    x = np.zeros((H, W), np.float32)
    x2 = np.zeros((H, W), np.float32)
    for s in range(SPP):
        rate = f(1) / f(1 + s)
        x = _lerp32(x, lum[..., s], rate)
        x2 = _lerp32(x2, lum[..., s] * lum[..., s], rate)
    var = np.abs(x2 - x * x)
    thr = float(np.median(var[var > 0]))
    _, histo, flags, _ = _serve(scene, thr, 1)
    assert 0 < flags.sum() < flags.size
    assert np.array_equal(flags.astype(bool), var > np.float32(thr))
    # the histogram of an unflagged pixel is its MRPNN samples'
    keep = flags == 0
    assert np.array_equal(histo[keep][:, 32], x[keep]) and np.array_equal(histo[keep][:, 31], x2[keep])


def test_fallback_pixels_are_delta_tracking_samples(scene):
    grid, md, mc, tc, lut, w, p = scene
    for k in (1, 3):
        rgb, histo, flags, lk = _serve(scene, 0.0, k)
        fl = flags.astype(bool)
        assert fl.sum() > 20 and lk > 0
        q = config1_params(size=(W, H), spp=k, multi=8, seed=7, spp_begin=SPP)
        q.light_color = p.light_color
        sums = np.zeros((H, W, 3), np.float64)
        oracle.render(grid, md, q, target=np.zeros((H, W, 3), np.float32), sum64=sums, threads=8)
        mean = sums / k
        if k == 1:
            assert np.array_equal(rgb[fl].astype(np.float64), mean[fl])
        else:
            np.testing.assert_allclose(rgb[fl], mean[fl], rtol=3e-7 * 4, atol=1e-9)
        assert np.all(histo[fl][:, 30] == k) and np.all(histo[~fl][:, 30] == SPP)
        # unflagged pixels keep the MRPNN estimate
        ref = oracle.render_mrpnn(grid, md, mc, tc, lut, w, p, threads=8)
        assert np.array_equal(rgb[~fl], ref[~fl])


def test_threshold_is_monotone(scene):
    prev = None
    for thr in (0.0, 1e-4, 1e-3, 1e-2, 1e-1):
        _, _, flags, _ = _serve(scene, thr, 1)
        if prev is not None:
            assert np.all(flags <= prev)
        prev = flags
