"""Throughput benchmark of the MRPNN rendering hot path on one MI355X.

Workload (N=1): the plain delta-tracking renderer on the 512^3 Perlin-fBm cloud
(seed 1, 5 octaves, base frequency 4, lacunarity 2, gain 0.5, threshold 0.4,
sphere falloff), extinction scale 256, albedo 0.999, HG g 0.857, sun
normalize(0,1,0.3) + sky 0.1, 1920x1080 pinhole (fov 40 deg, distance 2.5),
512 bounces.  One step = one 1080p frame of 1 sample per pixel folded into the
resident running mean (`VolumeRender::Render`, core/volume.cu:410-467, one
frame of its loop).  Inputs are resident in HBM before the timed region.

Prints one JSON line (stdout); progress goes to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "1080p frames/s, BF16 MLP vs plain delta tracking at equal RMSE; HBM/MFMA %"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LOOKUP = 32  # eight FP32 corners of one trilinear density lookup


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def config2_params(ia, width, height, spp_begin, spp_count, multi=512, seed=0):
    ori = 2.5 * np.array([0.75, 0.75, 0.75]) / np.linalg.norm([0.75, 0.75, 0.75])
    right, up = ia.camera_basis(ori, 40.0, width / height)
    light = np.array([0.0, 1.0, 0.3]) / np.linalg.norm([0.0, 1.0, 0.3])
    V = ia.Vec3.of
    return ia.RenderParams(width, height, V(ori), V(up), V(right), V(light), V((1, 1, 1)), V((0.1, 0.1, 0.1)),
                           V((0.999, 0.999, 0.999)), 0.857, 256.0, 1.0, multi, spp_begin, spp_count, seed,
                           ia.PT, 0)


def cpu_baseline(ia, vol, grid_dev, res, width, height, multi, target_s=15.0):
    """Time the CPU oracle (the reference algorithm restated in C++/OpenMP) on
    a bounded band of rows of the same frame; scale to frames/s."""
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    grid = grid_dev.download(np.float32, (res, res, res))
    md = vol.max_density
    p = config2_params(ia, width, height, 0, 1, multi)
    r0 = height // 2 - 2
    t0 = time.perf_counter()
    oracle.render(grid, md, p, threads=threads, rows=(r0, r0 + 4))
    probe = time.perf_counter() - t0
    rows = int(max(4, min(height, 4 * target_s / max(probe, 1e-3))))
    r0 = max(0, height // 2 - rows // 2)
    t0 = time.perf_counter()
    _, lookups, _ = oracle.render(grid, md, p, threads=threads, rows=(r0, r0 + rows))
    dt = time.perf_counter() - t0
    frames_per_s = (rows / height) / dt
    return {"value": frames_per_s, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "oracle tracer (C++/OpenMP, %d threads), rows %d..%d of one 1 spp 1920x1080 frame "
                      "(%.1f s, %.3g lookups/s), scaled to a full frame" % (threads, r0, r0 + rows, dt,
                                                                            lookups / dt)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bounces", type=int, default=512)
    ap.add_argument("--spp-per-step", type=int, default=1,
                    help="1-spp frames traced per launch (a step renders that many frames)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (reference kernel for HBM counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the image-tile / spp sharding is described in DESIGN.md)", file=sys.stderr)
        return 2

    import internship_amd as ia
    lib = ia.load()
    W, H, res = a.width, a.height, a.res
    if a.copy_calibration_mib > 0:
        nb = a.copy_calibration_mib << 20
        src, dst = ia.DeviceBuffer(nb), ia.DeviceBuffer(nb)
        src.zero()
        ia.check(lib.mrp_memcpy_d2d(dst.ptr, src.ptr, ctypes.c_size_t(nb)))
        log("calibration copy of %d MiB done" % a.copy_calibration_mib)
        src.free()
        dst.free()
    log("generating %d^3 cloud on the GPU" % res)
    grid_dev = ia.DeviceBuffer(res ** 3 * 4)
    ia.generate_cloud(res, seed=1, device_out=grid_dev)
    vol = ia.VolumeRender(res)
    vol.set_datas(grid_dev)
    vol.update()
    log("max density %.6f" % vol.max_density)
    target = ia.DeviceBuffer(W * H * 3 * 4)
    target.zero()
    st = ia.TraceStats()

    def step(i):
        p = config2_params(ia, W, H, i * a.spp_per_step, a.spp_per_step, a.bounces)
        vol.render_dev(p, target, None, st)
        return st.kernel_ms, st.lookups, st.paths, st.scatters

    for i in range(a.warmup):
        km, lk, _, _ = step(i)
        log("warmup %d: trace kernel %.2f ms, %.3g lookups" % (i, km, lk))
    kernel_ms = 0.0
    lookups = paths = scatters = 0
    t0 = time.perf_counter()
    for i in range(a.warmup, a.warmup + a.steps):
        km, lk, pa, sc = step(i)
        kernel_ms += km
        lookups += lk
        paths += pa
        scatters += sc
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / a.steps
    fps = a.steps * a.spp_per_step / elapsed
    avg_kernel_ms = kernel_ms / a.steps
    achieved = (lookups / a.steps) * BYTES_PER_LOOKUP / (avg_kernel_ms * 1e-3) / 1e9
    log("%.2f ms/frame, trace kernel %.2f ms, %.3g lookups/frame, %.1f GB/s algorithmic"
        % (ms_per_step, avg_kernel_ms, lookups / a.steps, achieved))
    img = target.download(np.float32, (H, W, 3))
    if not np.all(np.isfinite(img)):
        log("non-finite pixels in the image")
        return 1

    base = None
    if not a.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample")
        base = cpu_baseline(ia, vol, grid_dev, res, W, H, a.bounces)
        log("cpu baseline %.4g frames/s (%s)" % (base["value"], base["sample"]))

    out = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain delta-tracking tracer, %d^3 Perlin-fBm cloud (seed 1), %dx%d, 1 spp/frame, %d frames/launch, "
                               "%d bounces, alpha 256, albedo 0.999, g 0.857" % (res, W, H, a.spp_per_step, a.bounces),
                   "width": W, "height": H, "grid": res, "spp_per_step": a.spp_per_step, "bounces": a.bounces},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "trace_kernel", "kernel_ms": avg_kernel_ms,
                     "lookups_per_launch": lookups / a.steps, "bytes_per_lookup": BYTES_PER_LOOKUP},
        "cpu_baseline": base,
        "extra": {"paths_per_frame": paths / a.steps, "scatters_per_frame": scatters / a.steps,
                  "glookups_per_s": lookups / (kernel_ms * 1e-3) / 1e9, "max_path_steps": st.max_path_steps,
                  "image_mean": float(img.mean()), "grid_blocks": st.grid_blocks},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
