"""Training data of the learned part, generated on the GPU.

The reference's data flow (dataGen.cu:174-216,302) samples scattering points
in procedural clouds and renders `GetSample` targets there (core/render.cuh:
276-315); its descriptor is the RPNN one (dataGen.cu:53-105), so this build
trains on the MRPNN descriptor of the inference gather instead (B11).

Per cloud (one light, extinction scale, g and albedo, drawn per cloud):
  1. the Perlin-fBm cloud is generated on the GPU and prepared (`Update`);
  2. points are drawn in the box with probability proportional to density
     (host rejection sampling on the downloaded grid);
  3. a random frame (X, Y, Z) per point, X the viewing direction, as
     `GetMatrixFromNormal` draws them;
  4. descriptors: the gather kernel (mrp_mrpnn_features) -> 576 features + gamma;
  5. targets: mrp_get_samples (pos, X, light, alpha, g, albedo/1.001), spp
     paths each: `GetSample` multiplies 1.001 * scatter per bounce, and the
     renderer passes scatter_rate/1.001 to the network (core/render.cuh:498),
     so a per-bounce albedo of scatter_rate pairs with the input srp =
     (scatter_rate/1.001)^4.
A sample is (feat[576], aux = (gamma, g, srp, target)).

make_config4 builds the config-4 data set on the GPU (points from
`MeanFreePathSample`'s sampling, mrp_sample_training_points) straight into
HBM; sample_dataset and the loader of learn.py train on it.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import DeviceBuffer, VolumeRender, _lib, check, generate_cloud, srp_of
from ._lib import SampleQuery, Vec3


class DeviceView:
    """A pointer into a DeviceBuffer (for calls that write a slice)."""

    def __init__(self, buf, byte_offset):
        self.ptr = ctypes.c_void_p(buf.ptr.value + int(byte_offset))


@dataclass
class CloudSpec:
    seed: int
    threshold: float
    alpha: float
    g: float
    albedo: float
    light: tuple


def random_specs(n, seed, alpha=(64.0, 512.0), g=(0.7, 0.9), threshold=(0.3, 0.5), albedo=(0.9, 0.999)):
    """Config-4 style scenes: thresholds U[0.3,0.5], alpha U[64,512], g U[0.7,0.9],
    sun on the upper hemisphere."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = rng.normal(size=3)
        L[1] = abs(L[1]) + 0.2
        L /= np.linalg.norm(L)
        out.append(CloudSpec(seed + i, float(rng.uniform(*threshold)), float(rng.uniform(*alpha)),
                             float(rng.uniform(*g)), float(rng.uniform(*albedo)), tuple(float(x) for x in L)))
    return out


def sample_points(grid, n, rng):
    """n points in [-0.5, 0.5)^3 with probability proportional to the voxel density."""
    res = grid.shape[0]
    gmax = float(grid.max())
    if gmax <= 0:
        raise ValueError("empty cloud")
    pts = []
    need = n
    while need > 0:
        cand = rng.random((max(4 * need, 1024), 3))
        idx = np.minimum((cand * res).astype(np.int64), res - 1)
        d = grid[idx[:, 0], idx[:, 1], idx[:, 2]]
        keep = cand[rng.random(len(cand)) * gmax < d]
        pts.append(keep[:need])
        need -= len(pts[-1])
    return np.concatenate(pts) - 0.5


def random_frames(n, rng):
    X = rng.normal(size=(n, 3))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    r = rng.normal(size=(n, 3))
    Y = np.cross(X, r)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    Z = np.cross(X, Y)
    return X, Y, Z


def _f4(a, w):
    out = np.zeros((a.shape[0], 4), np.float32)
    out[:, :3] = a
    out[:, 3] = w
    return out


def make_dataset(specs, res=128, n_per_cloud=8192, spp=64, multi=128, sky=0.1, seed=0, log=None):
    """Returns (feat (N,576) FP32 host, aux (N,4) FP32, stats dict)."""
    import time
    lib = _lib.load()
    feats, auxs = [], []
    t_gather = t_target = 0.0
    lookups = 0
    rng = np.random.default_rng(seed)
    for ci, sp in enumerate(specs):
        gbuf = DeviceBuffer(res ** 3 * 4)
        generate_cloud(res, sp.seed, threshold=sp.threshold, device_out=gbuf)
        vol = VolumeRender(res)
        try:
            vol.set_datas(gbuf)
            vol.update()
            grid = gbuf.download(np.float32, (res, res, res))
            n = n_per_cloud
            pos = sample_points(grid, n, rng)
            X, Y, Z = random_frames(n, rng)
            bufs = []
            for arr in (_f4(pos, 1.0), _f4(X, 0.0), _f4(Y, 0.0), _f4(Z, 0.0)):
                b = DeviceBuffer(arr.nbytes)
                b.upload(arr)
                bufs.append(b)
            t0 = time.perf_counter()
            feat, gamma = vol.mrpnn_features(*bufs, n, sp.light, sp.alpha, sp.g, seed=0)  # the renderer's HG table
            t_gather += time.perf_counter() - t0
            q = (SampleQuery * n)()
            scatter = float(np.float32(sp.albedo) / np.float32(1.001))  # the network's albedo input, scatter/1.001
            for i in range(n):
                q[i] = SampleQuery(Vec3.of(pos[i]), Vec3.of(X[i]), Vec3.of(sp.light), sp.alpha, sp.g, scatter)
            qb = DeviceBuffer(ctypes.sizeof(q))
            check(lib.mrp_memcpy_h2d(qb.ptr, ctypes.cast(q, ctypes.c_void_p), ctypes.c_size_t(ctypes.sizeof(q))))
            tb = DeviceBuffer(n * 12)
            st = _lib.TraceStats()
            t0 = time.perf_counter()
            check(lib.mrp_get_samples(vol.h, n, qb.ptr, Vec3.of((1, 1, 1)), Vec3.of((sky, sky, sky)), multi, spp,
                                      ctypes.c_uint64(sp.seed), tb.ptr, ctypes.byref(st)))
            t_target += time.perf_counter() - t0
            lookups += st.lookups
            y = tb.download(np.float32, (n, 3))[:, 0]
            gam = gamma.download(np.float32, (n,))
            srp = srp_of((sp.albedo, sp.albedo, sp.albedo))[0]
            feats.append(feat.download(np.float32, (n, 576)))
            auxs.append(np.stack([gam, np.full(n, sp.g, np.float32), np.full(n, srp, np.float32), y], axis=1))
            if log:
                log("cloud %d/%d: seed %d, alpha %.0f, g %.3f, albedo %.3f, mean target %.4g"
                    % (ci + 1, len(specs), sp.seed, sp.alpha, sp.g, sp.albedo, float(y.mean())))
        finally:
            vol.close()
    feat = np.concatenate(feats)
    aux = np.concatenate(auxs).astype(np.float32)
    stats = {"gather_s": t_gather, "target_s": t_target, "target_lookups": lookups,
             "paths": len(specs) * n_per_cloud * spp}
    return feat, aux, stats


def relative_l2(pred, target):
    """sum (p - y)^2 / sum y^2."""
    pred = np.asarray(pred, np.float64)
    target = np.asarray(target, np.float64)
    return float(np.sum((pred - target) ** 2) / np.sum(target ** 2))


def predict(weights, feat, aux, precision=0):
    """The network (one channel) on host descriptors through mrp_mlp_forward:
    the samples are grouped by (g, srp), the forward's per-call scalars."""
    from . import Network
    net = Network(weights)
    out = np.zeros(len(feat), np.float32)
    try:
        keys = np.unique(aux[:, 1:3], axis=0)
        for g, srp in keys:
            sel = np.where((aux[:, 1] == g) & (aux[:, 2] == srp))[0]
            fb = DeviceBuffer(len(sel) * 576 * 4)
            fb.upload(np.ascontiguousarray(feat[sel]))
            gb = DeviceBuffer(len(sel) * 4)
            gb.upload(np.ascontiguousarray(aux[sel, 0]))
            o = net.forward(fb, gb, len(sel), float(g), (float(srp),) * 3, precision=precision)
            out[sel] = o.download(np.float32, (len(sel), 3))[:, 0]
    finally:
        net.close()
    return out


# --------------------------------------------------------------- config 4

@dataclass
class ViewSpec:
    cloud: int       # index of the cloud
    seed: int        # the cloud's Perlin seed
    threshold: float
    alpha: float
    g: float
    light: tuple


def config4_views(n_clouds=64, first_seed=100, views_per_cloud=4, seed=4):
    """Config 4's scenes: Perlin-fBm clouds with seeds first_seed.., threshold
    U[0.3,0.5] per cloud; per view (views_per_cloud per cloud) extinction scale
    U[64,512], a sun direction uniform on the upper hemisphere and HG g
    U[0.7,0.9]."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n_clouds):
        thr = float(rng.uniform(0.3, 0.5))
        for _ in range(views_per_cloud):
            L = rng.normal(size=3)
            L[1] = abs(L[1])
            L /= np.linalg.norm(L)
            out.append(ViewSpec(c, first_seed + c, thr, float(rng.uniform(64.0, 512.0)), float(rng.uniform(0.7, 0.9)),
                                tuple(float(x) for x in L)))
    return out


@dataclass
class Dataset:
    """Samples in HBM: feat (n x 576 FP32) and aux (n x 4: gamma, g, srp,
    target); cloud[i] is the cloud sample i comes from."""
    feat: DeviceBuffer
    aux: DeviceBuffer
    aux_host: np.ndarray
    cloud: np.ndarray
    n: int


def _frames_about(X, rng):
    """Random orthonormal frames (X, Y, Z) with the given first axes, uniform
    about them (`GetMatrixFromNormal`: Y = normalize(X x r), Z = Y x X)."""
    r = rng.normal(size=X.shape)
    Y = np.cross(X, r)
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    Z = np.cross(Y, X)
    return Y, Z


def make_config4(views, res=256, samples_per_view=4096, spp=256, bounces=512, albedo=0.999, sky=0.1, seed=1,
                 log=None):
    """Config-4 training data on the GPU, written straight into HBM.  Per view:
    the cloud's volume (built once per cloud), training points of
    `MeanFreePathSample`'s sampling, random frames about their view
    directions, the renderer's gather, and `GetSample` targets (spp paths of
    up to `bounces` scattering events, per-bounce albedo `albedo`)."""
    import time
    lib = _lib.load()
    n_total = len(views) * samples_per_view
    feat = DeviceBuffer(n_total * 576 * 4)
    aux = np.zeros((n_total, 4), np.float32)
    cloud = np.zeros(n_total, np.int32)
    rng = np.random.default_rng(seed)
    scatter = float(np.float32(albedo) / np.float32(1.001))
    srp = srp_of((albedo,) * 3)[0]
    stats = {"points_s": 0.0, "gather_s": 0.0, "target_s": 0.0, "target_lookups": 0, "paths": 0, "build_s": 0.0}
    vol = None
    cur = None
    off = 0
    t_all = time.perf_counter()
    try:
        for vi, v in enumerate(views):
            if cur != v.cloud:
                t0 = time.perf_counter()
                if vol is not None:
                    vol.close()
                gbuf = DeviceBuffer(res ** 3 * 4)
                generate_cloud(res, v.seed, threshold=v.threshold, device_out=gbuf)
                vol = VolumeRender(res)
                vol.set_datas(gbuf)
                vol.update()
                gbuf.free()
                cur = v.cloud
                stats["build_s"] += time.perf_counter() - t0
            m = samples_per_view
            # training points: rays until m points are kept
            t0 = time.perf_counter()
            pts, views_dir = [], []
            got, k = 0, 0
            while got < m:
                n_rays = max(1024, int(1.3 * (m - got)))
                pb = DeviceBuffer(n_rays * 64)
                vb = DeviceBuffer(n_rays * 64)
                check(lib.mrp_sample_training_points(vol.h, n_rays, ctypes.c_float(v.alpha), ctypes.c_float(v.g),
                                                     ctypes.c_uint64((seed << 40) + (vi << 8) + k), pb.ptr, vb.ptr,
                                                     None))
                p4 = pb.download(np.float32, (n_rays * 4, 4))
                v4 = vb.download(np.float32, (n_rays * 4, 4))
                keep = p4[:, 3] == 1.0
                pts.append(p4[keep, :3])
                views_dir.append(v4[keep, :3])
                got += int(keep.sum())
                k += 1
                if k > 64:
                    raise RuntimeError("view %d: too few scattering points (empty cloud?)" % vi)
            pos = np.concatenate(pts)[:m]
            X = np.concatenate(views_dir)[:m].astype(np.float64)
            stats["points_s"] += time.perf_counter() - t0
            Y, Z = _frames_about(X, rng)
            bufs = []
            for arr in (_f4(pos, 1.0), _f4(X, 0.0), _f4(Y, 0.0), _f4(Z, 0.0)):
                b = DeviceBuffer(arr.nbytes)
                b.upload(arr)
                bufs.append(b)
            t0 = time.perf_counter()
            gam = DeviceBuffer(m * 4)
            vol.mrpnn_features(*bufs, m, v.light, v.alpha, v.g, seed=0, feat=DeviceView(feat, off * 576 * 4),
                               gamma=gam)
            stats["gather_s"] += time.perf_counter() - t0
            q = np.zeros((m, 12), np.float32)
            q[:, 0:3] = pos
            q[:, 3:6] = X
            q[:, 6:9] = np.asarray(v.light, np.float32)
            q[:, 9] = v.alpha
            q[:, 10] = v.g
            q[:, 11] = scatter
            qb = DeviceBuffer(q.nbytes)
            qb.upload(q)
            tb = DeviceBuffer(m * 12)
            st = _lib.TraceStats()
            t0 = time.perf_counter()
            check(lib.mrp_get_samples(vol.h, m, qb.ptr, Vec3.of((1, 1, 1)), Vec3.of((sky, sky, sky)), bounces, spp,
                                      ctypes.c_uint64(seed * 1000003 + vi), tb.ptr, ctypes.byref(st)))
            stats["target_s"] += time.perf_counter() - t0
            stats["target_lookups"] += st.lookups
            stats["paths"] += m * spp
            y = tb.download(np.float32, (m, 3))[:, 0]
            aux[off:off + m] = np.stack([gam.download(np.float32, (m,)), np.full(m, v.g, np.float32),
                                         np.full(m, srp, np.float32), y], axis=1)
            cloud[off:off + m] = v.cloud
            off += m
            if log and (vi % max(1, len(views) // 16) == 0 or vi == len(views) - 1):
                log("view %d/%d: cloud seed %d, alpha %.0f, g %.3f, mean target %.4g (%.0f s so far)"
                    % (vi + 1, len(views), v.seed, v.alpha, v.g, float(y.mean()), time.perf_counter() - t_all))
    finally:
        if vol is not None:
            vol.close()
    auxb = DeviceBuffer(aux.nbytes)
    auxb.upload(aux)
    stats["total_s"] = time.perf_counter() - t_all
    return Dataset(feat, auxb, aux, cloud, n_total), stats
