"""MI355X-native MRPNN rendering hot path (Python side of the C ABI).

The native library (internship_amd/libmrp.so, built by
`python -m internship_amd.build`) implements everything; this module mirrors
the reference's host API so that tests and the benchmark read like the
reference's own usage:

  VolumeRender(resolution)                  core/volume.hpp:96
      .set_datas(grid | fill_func)          core/volume.cu:938-944 (SetData/SetDatas)
      .update()                             core/volume.cu:1203-1316
      .update_tr(light_dir, alpha)          core/volume.cu:1346-1424
      .update_hg_lut(g)                     core/volume.cu:1161-1183
      .set_scatter_rate/.set_tr_scale/...   core/volume.cu:1460-1493
      .render(size, ori, up, right, light_dir, rt, g, alpha, light_color,
              multi_scatter, sample_num)     core/volume.cu:410-467
  Camera(volume, name, position, resolution) core/camera.hpp:29
      .render_to_file(path, light_dir, ...)  core/camera.cu:226-252

There is no CPU fallback: every call goes to the HIP library and raises if
it is missing.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import (CloudParams, FrameParams, Histogram, MrpError, RenderParams, ServeParams, ServeStats, TraceStats,
                   Vec3, check, load)

PT = 0
MRPNN = 2

__all__ = ["VolumeRender", "Camera", "DeviceBuffer", "PT", "MRPNN", "MrpError", "camera_basis", "generate_cloud",
           "RenderParams", "TraceStats", "load"]


def _ptr(a: np.ndarray):
    # data_as keeps a reference to `a`, so temporaries made for the call stay alive through it
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """HBM allocation owned by Python (mrp_malloc / mrp_free)."""

    def __init__(self, nbytes: int):
        self.lib = load()
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(self.lib.mrp_malloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)))

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(self.lib.mrp_memcpy_h2d(self.ptr, _ptr(a), ctypes.c_size_t(a.nbytes)))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(self.lib.mrp_memcpy_d2h(_ptr(out), self.ptr, ctypes.c_size_t(out.nbytes)))
        return out

    def zero(self):
        check(self.lib.mrp_memset_zero(self.ptr, ctypes.c_size_t(self.nbytes)))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self.lib.mrp_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def camera_basis(ori, fov_deg=90.0, aspect=1.0):
    """Camera frame of `Camera::SetPosition` (core/camera.cu:173-189):
    forward = normalize(-ori), right = normalize(cross(Y, forward)),
    up = cross(forward, right); right/up are scaled by tan(fov/2) (vertical
    fov) and right additionally by the aspect ratio.  fov 90 = the reference."""
    ori = np.asarray(ori, dtype=np.float64)
    fwd = -ori / np.linalg.norm(ori)
    right = np.cross([0.0, 1.0, 0.0], fwd)
    n = np.linalg.norm(right)
    if n < 1e-12:
        right, up = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0 if ori[1] > 0 else -1.0])
    else:
        right = right / n
        up = np.cross(fwd, right)
    s = math.tan(math.radians(fov_deg) * 0.5)
    return (right * s * aspect).astype(np.float32), (up * s).astype(np.float32)


def read_hdr(path: str) -> np.ndarray:
    """A Radiance RGBE .hdr file as (h, w, 4) FP32 (mrp_read_hdr)."""
    lib = load()
    w, h = ctypes.c_int(), ctypes.c_int()
    check(lib.mrp_read_hdr(path.encode(), None, ctypes.c_size_t(0), ctypes.byref(w), ctypes.byref(h)))
    out = np.zeros((h.value, w.value, 4), np.float32)
    check(lib.mrp_read_hdr(path.encode(), _ptr(out), ctypes.c_size_t(out.shape[0] * out.shape[1]), ctypes.byref(w),
                           ctypes.byref(h)))
    return out


def generate_cloud(res, seed, threshold=0.4, octaves=5, base_freq=4.0, lacunarity=2.0, gain=0.5,
                   device_out=None) -> np.ndarray:
    """Perlin-fBm cloud on the GPU (host array, or into `device_out`)."""
    lib = load()
    p = CloudParams(seed, octaves, base_freq, lacunarity, gain, threshold)
    if device_out is not None:
        check(lib.mrp_generate_cloud(ctypes.byref(p), res, device_out.ptr, 1))
        return None
    out = np.empty((res, res, res), dtype=np.float32)
    check(lib.mrp_generate_cloud(ctypes.byref(p), res, _ptr(out), 0))
    return out


DATA_DIR = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "data")


def load_reference_weights() -> np.ndarray:
    """The trained MRPNN weights of core/NNWeight.cuh (49,673 FP32, header
    order), as committed in data/mrpnn_weights.bin; checked against the
    manifest's sha256."""
    import hashlib
    import json
    import os
    raw = open(os.path.join(DATA_DIR, "mrpnn_weights.bin"), "rb").read()
    man = json.load(open(os.path.join(DATA_DIR, "mrpnn_weights.json")))
    if hashlib.sha256(raw).hexdigest() != man["sha256"]:
        raise MrpError("mrpnn_weights.bin does not match its manifest checksum")
    return np.frombuffer(raw, dtype="<f4").copy()


def weight_manifest():
    import json
    import os
    return json.load(open(os.path.join(DATA_DIR, "mrpnn_weights.json")))


NATIVE_WEIGHTS = "mrpnn_weights_native"


def has_native_weights() -> bool:
    import os
    return os.path.exists(os.path.join(DATA_DIR, NATIVE_WEIGHTS + ".bin"))


def native_weights_manifest():
    import json
    import os
    return json.load(open(os.path.join(DATA_DIR, NATIVE_WEIGHTS + ".json")))


def load_native_weights() -> np.ndarray:
    """The MRPNN network trained by this library from seeded random weights on
    its own config-4 data (data/mrpnn_weights_native.bin, same layout as the
    shipped weights; the manifest holds the recipe, the command that produced
    the file and its held-out figures); checked against the manifest's sha256."""
    import hashlib
    import os
    raw = open(os.path.join(DATA_DIR, NATIVE_WEIGHTS + ".bin"), "rb").read()
    if hashlib.sha256(raw).hexdigest() != native_weights_manifest()["sha256"]:
        raise MrpError(NATIVE_WEIGHTS + ".bin does not match its manifest checksum")
    w = np.frombuffer(raw, dtype="<f4").copy()
    if w.size != 49673:
        raise MrpError(NATIVE_WEIGHTS + ".bin must hold 49,673 FP32 values")
    return w


class Network:
    """Trained MRPNN weights in HBM (FP32 + BF16 MFMA packing)."""

    def __init__(self, weights: np.ndarray = None):
        self.lib = load()
        w = load_reference_weights() if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self.weights = w
        self.h = ctypes.c_void_p()
        check(self.lib.mrp_network_create(_ptr(w), ctypes.c_size_t(w.size), ctypes.byref(self.h)))

    def forward(self, feat: "DeviceBuffer", gamma: "DeviceBuffer", n: int, g: float, srp, precision=0,
                out: "DeviceBuffer" = None) -> "DeviceBuffer":
        out = out or DeviceBuffer(max(1, n) * 12)
        check(self.lib.mrp_mlp_forward(self.h, int(n), feat.ptr, gamma.ptr, float(g), Vec3.of(srp), int(precision),
                                       out.ptr))
        return out

    def forward_tiled(self, tiles: "DeviceBuffer", pool: "DeviceBuffer", gamma: "DeviceBuffer", n: int, g: float,
                      srp, precision=1, out: "DeviceBuffer" = None) -> "DeviceBuffer":
        out = out or DeviceBuffer(max(1, n) * 12)
        check(self.lib.mrp_mlp_forward_tiled(self.h, int(n), tiles.ptr, pool.ptr, gamma.ptr, float(g), Vec3.of(srp),
                                             int(precision), out.ptr))
        return out

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.mrp_network_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trainer:
    """GPU training of the MRPNN weights (mrp_trainer_*): forward and backward
    in FP32 (precision 0) or on BF16 matrix cores with FP32 accumulation
    (precision 1), FP64 weight gradients, Adam on FP32 master weights.  A batch
    is feat (B x 576 FP32 in HBM, the gather's layout) and aux (B x 4 FP32:
    gamma, g, srp, target)."""

    def __init__(self, weights: np.ndarray = None, max_batch: int = 16384, precision: int = 0):
        self.lib = load()
        w = load_reference_weights() if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self.h = ctypes.c_void_p()
        self.max_batch = int(max_batch)
        check(self.lib.mrp_trainer_create(_ptr(w), ctypes.c_size_t(w.size), self.max_batch, ctypes.byref(self.h)))
        if precision:
            self.set_precision(precision)

    def set_precision(self, precision: int):
        check(self.lib.mrp_trainer_set_precision(self.h, int(precision)))

    def gradient(self, feat: "DeviceBuffer", aux: "DeviceBuffer", batch: int):
        g = np.zeros(49673, dtype=np.float64)
        loss = ctypes.c_double()
        check(self.lib.mrp_trainer_gradient(self.h, int(batch), feat.ptr, aux.ptr, _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def step(self, feat: "DeviceBuffer", aux: "DeviceBuffer", batch: int, lr: float = 1e-3) -> float:
        loss = ctypes.c_double()
        check(self.lib.mrp_trainer_step(self.h, int(batch), feat.ptr, aux.ptr, float(lr), ctypes.byref(loss)))
        return loss.value

    def weights(self) -> np.ndarray:
        out = np.zeros(49673, dtype=np.float32)
        check(self.lib.mrp_trainer_get_weights(self.h, _ptr(out)))
        return out

    def predictions(self, batch: int) -> np.ndarray:
        out = np.zeros(int(batch), dtype=np.float32)
        check(self.lib.mrp_trainer_predictions(self.h, int(batch), _ptr(out)))
        return out

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.mrp_trainer_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TILE_BYTES = 3 * 12 * 64 * 16  # one tile: 32 descriptors x 576 BF16 features


def untile_descriptors(raw: np.ndarray, n: int) -> np.ndarray:
    """Wave-tiled BF16 descriptors (mrp_mrpnn_features_tiled) -> (n, 576) float32.
    Chunk (tile t, stream k, chunk c, lane 32 kg + d) holds the 8 BF16 features
    16c + 8h + 4kg + e (h in 0..1, e in 0..3) of descriptor 32 t + d."""
    t = (n + 31) // 32
    u16 = np.frombuffer(np.ascontiguousarray(raw).tobytes()[:t * TILE_BYTES], dtype=np.uint16)
    ch = u16.reshape(t, 3, 12, 2, 32, 2, 4)  # tile, stream, chunk, kg, d, h, e
    f = ch.transpose(0, 4, 1, 2, 5, 3, 6)  # tile, d, stream, chunk, h, kg, e
    bits = f.reshape(t * 32, 576)[:n].astype(np.uint32) << 16
    return bits.view(np.float32)


def srp_of(scatter_rate):
    """(scatter_rate / 1.001)^4 per channel, as `RadiancePredict` forms it."""
    sr = np.asarray(scatter_rate, dtype=np.float32) / np.float32(1.001)
    return tuple(float(np.float32(float(x) ** 4.0)) for x in sr)


class VolumeRender:
    """Handle-based twin of the reference `VolumeRender` (core/volume.hpp:45-146)."""

    PT = PT
    MRPNN = MRPNN

    def __init__(self, resolution: int):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.mrp_volume_create(int(resolution), ctypes.byref(self.h)))
        self.resolution = int(resolution)
        # state of the reference's device-global setters (core/render.cuh:14-16)
        self.scatter_rate = (1.001, 1.001, 1.001)
        self.tr_scale = 1.0
        self.sky = (0.1, 0.1, 0.1)  # constant sky (the reference needs an HDRI, core/render.cuh:57-59)
        self.exposure = 1.0
        self.checkerboard = True  # the reference's default (core/volume.hpp:50)
        self.seed = 0

    # ---- data
    def set_datas(self, data):
        """`SetDatas(FillFunc)`: `data` is a (res,res,res) array indexed [x,y,z],
        a callable f(x,y,z,u,v,w) evaluated on numpy index grids, or a
        DeviceBuffer holding res^3 FP32 in HBM."""
        r = self.resolution
        if isinstance(data, DeviceBuffer):
            check(self.lib.mrp_volume_set_data(self.h, data.ptr, 1))
            return
        if callable(data):
            i = np.arange(r)
            x, y, z = np.meshgrid(i, i, i, indexing="ij")
            u, v, w = ((x + 0.5) / r), ((y + 0.5) / r), ((z + 0.5) / r)
            data = np.asarray(data(x, y, z, u.astype(np.float32), v.astype(np.float32), w.astype(np.float32)))
        a = np.ascontiguousarray(data, dtype=np.float32)
        if a.shape != (r, r, r):
            raise ValueError("grid must be (%d,%d,%d)" % (r, r, r))
        check(self.lib.mrp_volume_set_data(self.h, _ptr(a), 0))

    def update(self):
        check(self.lib.mrp_volume_update(self.h))

    @property
    def max_density(self) -> float:
        f = ctypes.c_float()
        check(self.lib.mrp_volume_max_density(self.h, ctypes.byref(f)))
        return f.value

    def get_mip(self, m: int) -> np.ndarray:
        r = 256 >> m
        out = np.empty((r, r, r), dtype=np.float32)
        check(self.lib.mrp_volume_get_mip(self.h, m, _ptr(out)))
        return out

    def get_occupancy(self):
        """The trackers' brick occupancy: (bool array n x n x n over bricks of
        2^shift cells per axis, shift), see mrp_volume_get_occupancy."""
        sh, n = ctypes.c_int(), ctypes.c_int()
        check(self.lib.mrp_volume_get_occupancy(self.h, None, 0, ctypes.byref(sh), ctypes.byref(n)))
        nb = n.value ** 3
        words = np.zeros((nb + 31) // 32, np.uint32)
        check(self.lib.mrp_volume_get_occupancy(self.h, _ptr(words), words.size, ctypes.byref(sh), ctypes.byref(n)))
        bits = (words[:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & 1
        return bits.ravel()[:nb].astype(bool).reshape(n.value, n.value, n.value), sh.value

    def update_tr(self, light_dir, alpha=64.0):
        check(self.lib.mrp_volume_update_tr(self.h, Vec3.of(light_dir), float(alpha)))

    def get_tr_mip(self, m: int) -> np.ndarray:
        r = 128 >> m
        out = np.empty((r, r, r), dtype=np.float32)
        check(self.lib.mrp_volume_get_tr_mip(self.h, m, _ptr(out)))
        return out

    def update_hg_lut(self, g, seed=0):
        check(self.lib.mrp_volume_update_hg_lut(self.h, float(g), ctypes.c_uint64(seed)))

    def get_hg_lut(self) -> np.ndarray:
        out = np.empty((512, 512), dtype=np.float32)
        check(self.lib.mrp_volume_get_hg_lut(self.h, _ptr(out)))
        return out

    # ---- setters (core/volume.cu:1460-1493)
    def set_scatter_rate(self, rate):
        self.scatter_rate = tuple(float(x) for x in (rate if np.ndim(rate) else (rate, rate, rate)))

    def set_tr_scale(self, s):
        self.tr_scale = float(s)

    def set_env(self, sky):
        """Constant environment radiance (stands in for SetHDRI/SetEnvExp)."""
        self.sky = tuple(float(x) for x in (sky if np.ndim(sky) else (sky, sky, sky)))

    def set_exposure(self, e):
        self.exposure = float(e)

    def set_checkerboard(self, on: bool):
        self.checkerboard = bool(on)

    # ---- rendering
    def params(self, size, ori, up, right, light_dir, rt=PT, g=0.857, alpha=1.0, light_color=(1, 1, 1),
               multi_scatter=512, sample_num=1024, spp_begin=0, seed=None, mlp_precision=0) -> RenderParams:
        w, h = (size, size) if np.isscalar(size) else (int(size[0]), int(size[1]))
        return RenderParams(w, h, Vec3.of(ori), Vec3.of(up), Vec3.of(right), Vec3.of(light_dir),
                            Vec3.of(light_color), Vec3.of(self.sky), Vec3.of(self.scatter_rate), float(g),
                            float(alpha), float(self.tr_scale), int(multi_scatter), int(spp_begin),
                            int(sample_num), int(self.seed if seed is None else seed), int(rt), int(mlp_precision))

    def render(self, size, ori, up, right, light_dir, rt=PT, g=0.857, alpha=1.0, light_color=(1, 1, 1),
               multi_scatter=512, sample_num=1024, seed=None, stats: TraceStats = None,
               mlp_precision=0) -> np.ndarray:
        """Offline render (`VolumeRender::Render`, core/volume.cu:410-467):
        returns (height, width, 3) FP32, the mean of `sample_num` samples.
        mlp_precision (MRPNN): 0 FP32, 1 BF16 MFMA, 2 FP8-E4M3 MFMA."""
        p = self.params(size, ori, up, right, light_dir, rt, g, alpha, light_color, multi_scatter, sample_num,
                        0, seed, mlp_precision)
        out = np.zeros((p.height, p.width, 3), dtype=np.float32)
        check(self.lib.mrp_render_offline(self.h, ctypes.byref(p), _ptr(out),
                                          ctypes.byref(stats) if stats is not None else None))
        return out

    def render_dev(self, p: RenderParams, target: DeviceBuffer, sum64: DeviceBuffer = None,
                   stats: TraceStats = None):
        check(self.lib.mrp_render_offline_dev(self.h, ctypes.byref(p), target.ptr,
                                              sum64.ptr if sum64 is not None else None,
                                              ctypes.byref(stats) if stats is not None else None))

    def render_frame(self, target: DeviceBuffer, histo: DeviceBuffer, target2: DeviceBuffer, size, ori, up, right,
                     light_dir, light_color=(1, 1, 1), alpha=1.0, multi_scatter=1, g=0.0, randseed=0, rt=PT,
                     tone=2, denoise=False, mlp_precision=1, stats: TraceStats = None):
        """One progressive frame (`VolumeRender::Render(float3* target, Histogram*
        histo, unsigned* target2, ...)`, core/volume.cu:469-547) into
        caller-owned HBM buffers: target (w*h*3 FP32), histo (w*h Histogram),
        target2 (w*h BGRA8 words)."""
        p = self.params(size, ori, up, right, light_dir, rt, g, alpha, light_color, multi_scatter, 1, randseed,
                        None, mlp_precision)
        fp = FrameParams(p, int(tone), 1 if denoise else 0, 1 if self.checkerboard else 0, float(self.exposure))
        check(self.lib.mrp_render_frame(self.h, ctypes.byref(fp), target.ptr, histo.ptr, target2.ptr,
                                        ctypes.byref(stats) if stats is not None else None))

    def reset_frames(self):
        check(self.lib.mrp_volume_reset_frames(self.h))

    def frame_state(self):
        flip, calls = ctypes.c_int(), ctypes.c_uint32()
        check(self.lib.mrp_volume_frame_state(self.h, ctypes.byref(flip), ctypes.byref(calls)))
        return flip.value, calls.value

    def set_network(self, net: "Network"):
        self.network = net
        check(self.lib.mrp_volume_set_network(self.h, net.h if net is not None else None))

    def mrpnn_features(self, pos_act: DeviceBuffer, fx: DeviceBuffer, fy: DeviceBuffer, fz: DeviceBuffer, n: int,
                       light_dir, alpha, g, seed=0, feat: DeviceBuffer = None, gamma: DeviceBuffer = None):
        feat = feat or DeviceBuffer(max(1, n) * 576 * 4)
        gamma = gamma or DeviceBuffer(max(1, n) * 4)
        check(self.lib.mrp_mrpnn_features(self.h, int(n), pos_act.ptr, fx.ptr, fy.ptr, fz.ptr, Vec3.of(light_dir),
                                          float(alpha), float(g), ctypes.c_uint64(seed), feat.ptr, gamma.ptr))
        return feat, gamma

    def mrpnn_features_tiled(self, pos_act: DeviceBuffer, fx: DeviceBuffer, fy: DeviceBuffer, fz: DeviceBuffer,
                             n: int, light_dir, alpha, g, seed=0):
        """The BF16/FP8 renderer's gather: (tiles, pool, gamma) device buffers;
        decode the tiles with `untile_descriptors`."""
        tiles = DeviceBuffer(max(1, (n + 31) // 32) * TILE_BYTES)
        pool = DeviceBuffer(max(1, n) * 72 * 4)
        gamma = DeviceBuffer(max(1, n) * 4)
        check(self.lib.mrp_mrpnn_features_tiled(self.h, int(n), pos_act.ptr, fx.ptr, fy.ptr, fz.ptr,
                                                Vec3.of(light_dir), float(alpha), float(g), ctypes.c_uint64(seed),
                                                tiles.ptr, pool.ptr, gamma.ptr))
        return tiles, pool, gamma

    def set_hdri(self, rgba, exposure: float = 1.0):
        """Environment map (rgba (h, w, 4) FP32, row 0 at the top) for `SkyBox`, or None to use the constant sky."""
        if rgba is None:
            check(self.lib.mrp_volume_set_hdri(self.h, None, 0, 0, ctypes.c_float(1.0)))
            return
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        check(self.lib.mrp_volume_set_hdri(self.h, _ptr(a), int(a.shape[1]), int(a.shape[0]),
                                           ctypes.c_float(exposure)))

    def set_gather_cells(self, bits: int):
        """32 (FP32 cells, default) or 16 (FP16 cells) for the BF16/FP8 renderers' gather."""
        check(self.lib.mrp_volume_set_gather_cells(self.h, int(bits)))

    def set_sample_buffer_budget(self, nbytes: int):
        check(self.lib.mrp_set_sample_buffer_budget(self.h, ctypes.c_size_t(int(nbytes))))

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.mrp_volume_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeBatch:
    """Output buffers of mrp_serve_views for n_views views of width x height
    (HBM): rgb (FP32 estimate before the denoise), bgra (packed output) and
    the fallback flags."""

    def __init__(self, n_views: int, width: int, height: int):
        self.n_views, self.width, self.height = int(n_views), int(width), int(height)
        n = self.n_views * self.width * self.height
        self.rgb = DeviceBuffer(n * 12)
        self.bgra = DeviceBuffer(n * 4)
        self.flags = DeviceBuffer(n)

    def run(self, volumes, views, mrpnn_spp=2, var_threshold=0.01, fallback_spp=16, denoise=True, tone=2,
            exposure=1.0, bgra=True, flags=True) -> ServeStats:
        """One batched call: views[k] (RenderParams) on volumes[k] (VolumeRender)."""
        lib = load()
        n = len(views)
        assert n == len(volumes) == self.n_views
        hs = (ctypes.c_void_p * n)(*[v.h.value for v in volumes])
        ps = (RenderParams * n)(*views)
        sp = ServeParams(int(mrpnn_spp), float(var_threshold), int(fallback_spp), int(bool(denoise)), int(tone),
                         float(exposure))
        st = ServeStats()
        check(lib.mrp_serve_views(ctypes.cast(hs, ctypes.c_void_p), ctypes.cast(ps, ctypes.c_void_p), n,
                                  ctypes.byref(sp), self.rgb.ptr, self.bgra.ptr if bgra else None,
                                  self.flags.ptr if flags else None, ctypes.byref(st)))
        return st

    def download(self):
        """(rgb (n, h, w, 3) FP32, bgra (n, h, w) uint32, flags (n, h, w) uint8)."""
        shp = (self.n_views, self.height, self.width)
        return (self.rgb.download(np.float32, shp + (3,)), self.bgra.download(np.uint32, shp),
                self.flags.download(np.uint8, shp))


def aces(c: np.ndarray) -> np.ndarray:
    """Host ACES tone map of `Camera::ACES` (core/camera.cu:74-160) - used by
    `Camera.render_to_file`."""
    from .tonemap import aces_np
    return aces_np(c)


class Camera:
    """`Camera` (core/camera.hpp:10-41) over a VolumeRender."""

    def __init__(self, volume: VolumeRender, name="test camera", position=(0.75, 0.75, 0.75), resolution=512,
                 fov_deg=90.0):
        self.volume = volume
        self.name = name
        self.resolution = resolution
        self.fov_deg = fov_deg
        self.set_position(position)

    def set_position(self, position):
        self.ori = np.asarray(position, dtype=np.float32)
        self.right, self.up = camera_basis(self.ori, self.fov_deg)

    def render(self, light_dir, light_color=(1, 1, 1), alpha=1.0, multi_scatter=1, g=0.0, sample_num=1,
               rt=PT) -> np.ndarray:
        return self.volume.render(self.resolution, self.ori, self.up, self.right, light_dir, rt, g, alpha,
                                  light_color, multi_scatter, sample_num)

    def render_to_file(self, path, light_dir, light_color=(1, 1, 1), alpha=1.0, multi_scatter=1, g=0.0,
                       sample_num=1, tone="aces", rt=PT, exp=1.0):
        """`Camera::RenderToFile` (core/camera.cu:226-252): render, tone map,
        write `path.bmp` (32-bit BGRA, top-down)."""
        from .tonemap import save_bmp, tone_map
        img = self.render(light_dir, light_color, alpha, multi_scatter, g, sample_num, rt)
        save_bmp(tone_map(img * exp, tone), path + ".bmp")
        return img
