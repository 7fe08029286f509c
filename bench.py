"""Throughput benchmark of the MRPNN rendering hot path on one MI355X.

Scene (N=1): the 512^3 Perlin-fBm cloud (seed 1, 5 octaves, base frequency 4,
lacunarity 2, gain 0.5, threshold 0.4, sphere falloff), extinction scale 256,
albedo 0.999, HG g 0.857, sun normalize(0,1,0.3) + constant sky 0.1,
1920x1080 pinhole (vertical fov 40 deg, distance 2.5), 512 bounces.
Inputs are generated on the GPU and resident in HBM before any timed region.

Default mode (headline): frames/s of the BF16 MLP renderer at 1080p, one
sample per pixel per frame (`VolumeRender::Render` with RenderType MRPNN,
core/volume.cu:410-467), timed over K frames after W warm-up frames.  Equal-
RMSE protocol beside it: RMSE over linear RGB against a 4096-spp delta-tracking
ground truth with FP64 per-pixel sums; the smallest delta-tracking spp N with
RMSE(DT_N) <= RMSE(MLP) (powers of two, then bisection); speed-up =
fps(MLP) / fps(DT_N), where fps(DT_N) is measured on one N-spp launch.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode headline|mlp|tracer|frame|train|serve]
                  [--no-cpu-baseline] [--truth-spp 4096]

Mode mlp: only the timed MLP-renderer frames of the headline (for profiling).
Mode frame: progressive path-traced frames with histograms, denoiser and tone map.
Mode train: GPU training data (descriptors + GetSample targets) and Adam steps, with an A/B test.
Mode serve: 16 scenes x 4 cameras at 720p served by concurrent host threads; p50/p99 latency.
Mode tracer: frames/s of the plain tracer (1 spp per frame, --spp-per-step
frames per launch) with its HBM roofline.  Prints one JSON line on stdout;
progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "1080p frames/s, BF16 MLP vs plain delta tracking at equal RMSE; HBM/MFMA %"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0
FP32_PEAK_TFLOPS = 157.3
BYTES_PER_LOOKUP = 32          # eight FP32 corners of one trilinear density lookup
MLP_FLOP_PER_PASS = 97072      # 48,536 MACs per descriptor per colour pass
GATHER_BYTES_PER_POINT = 18400  # 192 x 32 (density) + 192 x 32 (transmittance) + 382 x 16 (HG table)
GATHER_BYTES_PER_POINT_16 = 12256  # the same with 16-byte FP16 density and transmittance cells
FRAME_POST_BYTES_PER_PIXEL = 12 + 132 + 132 + 12 + 4  # update: target + histogram r/w; denoise: read both; pack
MLP_WORKLOAD = "cloud512_1080p_1spp"
TRACER_WORKLOAD = "cloud512_1080p_%dspp_per_launch"

TRUTH_SEED = 1000
DT_SEED = 2000
MLP_SEED = 0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def scene_params(ia, width, height, spp_begin, spp_count, multi=512, seed=0, rt=0, precision=1):
    ori = 2.5 * np.array([0.75, 0.75, 0.75]) / np.linalg.norm([0.75, 0.75, 0.75])
    right, up = ia.camera_basis(ori, 40.0, width / height)
    light = np.array([0.0, 1.0, 0.3]) / np.linalg.norm([0.0, 1.0, 0.3])
    V = ia.Vec3.of
    return ia.RenderParams(width, height, V(ori), V(up), V(right), V(light), V((1, 1, 1)), V((0.1, 0.1, 0.1)),
                           V((0.999, 0.999, 0.999)), 0.857, 256.0, 1.0, multi, spp_begin, spp_count, seed,
                           rt, precision)


def rmse(img, ref):
    d = img.astype(np.float64) - ref
    return float(np.sqrt(np.mean(d * d)))


def luma_rmse100(img, ref):
    """RMSE x 100 of tone-mapped 8-bit luma, the reference's image metric
    (tools/Compare.py:31-42), with ACES (core/camera.cu:74-160)."""
    from internship_amd.tonemap import aces_np

    def luma8(x):
        t = np.clip(aces_np(np.asarray(x, np.float32)), 0, 1)
        v = np.floor(t * 255.0)
        return (0.299 * v[..., 0] + 0.587 * v[..., 1] + 0.114 * v[..., 2]) / 255.0

    d = luma8(img) - luma8(ref)
    return float(100.0 * np.sqrt(np.mean(d * d)))


def luma_rmse100_packed(bgra, ref):
    """luma_rmse100 of a packed 0xAARRGGBB image (ACES, exposure 1) against a linear reference."""
    from internship_amd.tonemap import aces_np
    b = np.asarray(bgra, np.uint32)
    v = np.stack([(b >> 16) & 255, (b >> 8) & 255, b & 255], axis=-1).astype(np.float64)
    la = (0.299 * v[..., 0] + 0.587 * v[..., 1] + 0.114 * v[..., 2]) / 255.0
    t = np.floor(np.clip(aces_np(np.asarray(ref, np.float32)), 0, 1) * 255.0)
    lb = (0.299 * t[..., 0] + 0.587 * t[..., 1] + 0.114 * t[..., 2]) / 255.0
    return float(100.0 * np.sqrt(np.mean((la - lb) ** 2)))


class Scene:
    def __init__(self, ia, res, width, height, bounces):
        self.ia, self.res, self.W, self.H, self.bounces = ia, res, width, height, bounces
        log("generating %d^3 cloud on the GPU" % res)
        self.grid = ia.DeviceBuffer(res ** 3 * 4)
        ia.generate_cloud(res, seed=1, device_out=self.grid)
        self.vol = ia.VolumeRender(res)
        self.vol.set_datas(self.grid)
        self.vol.update()
        log("max density %.6f" % self.vol.max_density)
        self.target = ia.DeviceBuffer(width * height * 3 * 4)
        self.st = ia.TraceStats()
        self._params = {}

    def render(self, spp_begin, spp_count, seed, rt=0, precision=1, sum64=None, fresh=True):
        if fresh:
            self.target.zero()
        # the parameters of a frame differ only in the first sample index: build them once per kind
        key = (spp_count, seed, rt, precision)
        p = self._params.get(key)
        if p is None:
            p = self._params[key] = scene_params(self.ia, self.W, self.H, spp_begin, spp_count, self.bounces, seed, rt,
                                                 precision)
        p.spp_begin = spp_begin
        t0 = time.perf_counter()
        self.vol.render_dev(p, self.target, sum64, self.st)
        return time.perf_counter() - t0

    def image(self):
        return self.target.download(np.float32, (self.H, self.W, 3))


def ground_truth_image(ia, sc, spp):
    """Ground-truth image: `spp` frames of the GPU tracer summed in FP64 per pixel."""
    sum64 = ia.DeviceBuffer(sc.W * sc.H * 3 * 8)
    sum64.zero()
    batch = 64
    t0 = time.perf_counter()
    for s in range(0, spp, batch):
        n = min(batch, spp - s)
        sc.render(s, n, TRUTH_SEED, sum64=sum64, fresh=(s == 0))
        if (s // batch) % 16 == 0:
            log("ground truth: %d / %d spp (%.1f s)" % (s + n, spp, time.perf_counter() - t0))
    img = sum64.download(np.float64, (sc.H, sc.W, 3)) / spp
    log("ground truth %d spp done in %.1f s, mean %.5f" % (spp, time.perf_counter() - t0, img.mean()))
    return img


def dt_rmse(sc, n, ref):
    sc.render(0, n, DT_SEED)
    return rmse(sc.image(), ref)


def dt_time(sc, n, reps=2):
    best = (1e30, 0.0, 0)
    for _ in range(reps):
        t = sc.render(0, n, DT_SEED)
        if t < best[0]:
            best = (t, sc.st.kernel_ms, sc.st.lookups)
    return best


def equal_rmse_search(sc, target_rmse, ref, max_spp):
    """Smallest N with RMSE(DT_N) <= target: powers of two, then bisection."""
    hist = {}
    n = 1
    while True:
        hist[n] = dt_rmse(sc, n, ref)
        log("DT %d spp: RMSE %.5g (target %.5g)" % (n, hist[n], target_rmse))
        if hist[n] <= target_rmse or n >= max_spp:
            break
        n *= 2
    if hist[n] > target_rmse:
        return n, hist, False
    lo, hi = n // 2, n
    while hi - lo > 1 and lo >= 1:
        mid = (lo + hi) // 2
        hist[mid] = dt_rmse(sc, mid, ref)
        log("DT %d spp: RMSE %.5g" % (mid, hist[mid]))
        if hist[mid] <= target_rmse:
            hi = mid
        else:
            lo = mid
    return hi, hist, True


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, target_s=15.0, single_s=8.0):
    """Time the CPU oracle (the reference tracer restated in C++/OpenMP) on a
    bounded sample of the same scene: whole 1-spp frames, as many as fit the
    target time; frames/s.  Threads: every core the process may run on
    (sched_getaffinity), bounded by OMP_NUM_THREADS when the job sets it (the
    GPU box gives a one-GPU job 16 cores).  Beside it a single-thread figure
    from a band of rows: its lookup rate over the lookups of a whole frame."""
    import oracle
    from internship_amd.build import GXX_FLAGS
    visible = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(visible, omp) if omp > 0 else visible
    grid = sc.grid.download(np.float32, (sc.res, sc.res, sc.res))
    md = sc.vol.max_density

    def frames(nthreads, budget_s):
        # whole 1-spp frames: one to size the sample, then enough for the budget
        t0 = time.perf_counter()
        _, lk0, _ = oracle.render(grid, md, scene_params(sc.ia, sc.W, sc.H, 0, 1, sc.bounces, DT_SEED),
                                  threads=nthreads)
        t1 = time.perf_counter() - t0
        n = int(max(0, min(63, round(budget_s / max(t1, 1e-3)) - 1)))
        lk = 0
        if n:
            t0 = time.perf_counter()
            _, lk, _ = oracle.render(grid, md, scene_params(sc.ia, sc.W, sc.H, 1, n, sc.bounces, DT_SEED),
                                     threads=nthreads)
            t1 += time.perf_counter() - t0
        return (n + 1) / t1, n + 1, t1, lk0 + lk

    def band(nthreads, budget_s):
        # rows through the middle of the cloud, sized for the budget
        r0 = sc.H // 2 - 2
        t0 = time.perf_counter()
        oracle.render(grid, md, scene_params(sc.ia, sc.W, sc.H, 0, 1, sc.bounces, DT_SEED), threads=nthreads,
                      rows=(r0, r0 + 4))
        probe = time.perf_counter() - t0
        rows = int(max(4, min(sc.H, 4 * budget_s / max(probe, 1e-3))))
        r0 = max(0, sc.H // 2 - rows // 2)
        t0 = time.perf_counter()
        _, lookups, _ = oracle.render(grid, md, scene_params(sc.ia, sc.W, sc.H, 0, 1, sc.bounces, DT_SEED),
                                      threads=nthreads, rows=(r0, r0 + rows))
        return r0, rows, time.perf_counter() - t0, lookups

    fps, nf, dt, lookups = frames(threads, target_s)
    lk_per_frame = lookups / nf
    r01, rows1, dt1, lookups1 = band(1, single_s)
    fps1 = lookups1 / dt1 / lk_per_frame
    return {"value": fps, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "plain delta-tracking tracer, CPU oracle (C++/OpenMP, %d threads), %d whole 1-spp %dx%d "
                      "frames, %d bounces (%.1f s, %.3g lookups/s)"
                      % (threads, nf, sc.W, sc.H, sc.bounces, dt, lookups / dt),
            "single_thread": {"value": fps1, "unit": "frames/s",
                              "sample": "1 thread, rows %d..%d of one frame (%.1f s, %.3g lookups/s); frames/s = "
                                        "that lookup rate over the %.3g lookups of a whole frame"
                                        % (r01, r01 + rows1, dt1, lookups1 / dt1, lk_per_frame)},
            "cores_visible": visible, "omp_num_threads": omp or None, "cpu_model": cpu_model(),
            "compiler": "g++ " + " ".join(f for f in GXX_FLAGS if not f.startswith("-I"))}


CEILING_BUFFER = 4 << 30  # about the size of the 512^3 cell array (4.3 GB)


def random_read_ceiling(ia):
    """The GPU's rate for random 32-byte reads (the access shape of one
    trilinear cell lookup) over a 4 GiB buffer: mrp_measure_random_reads."""
    r = ctypes.c_double(0.0)
    ia.check(ia.load().mrp_measure_random_reads(BYTES_PER_LOOKUP, ctypes.c_uint64(CEILING_BUFFER), 64, ctypes.byref(r)))
    log("random 32-byte reads over 4 GiB: %.1f G/s (%.0f GB/s useful, %.0f GB/s of 128-byte lines)"
        % (r.value / 1e9, r.value * BYTES_PER_LOOKUP / 1e9, r.value * 128 / 1e9))
    return {"width_bytes": BYTES_PER_LOOKUP, "buffer_bytes": CEILING_BUFFER, "reads_per_s": r.value,
            "useful_gbs": r.value * BYTES_PER_LOOKUP / 1e9}


def with_ceiling(roof, ceiling, lookups_per_s):
    roof = dict(roof)
    roof["frac_of_random_read_ceiling"] = lookups_per_s / ceiling["reads_per_s"]
    return roof


def run_tracer(a, ia, sc):
    def step(i):
        p = scene_params(ia, sc.W, sc.H, i * a.spp_per_step, a.spp_per_step, sc.bounces, 0, 0, 0)
        sc.vol.render_dev(p, sc.target, None, sc.st)
        return sc.st.kernel_ms, sc.st.lookups

    sc.target.zero()
    for i in range(a.warmup):
        step(i)
    kernel_ms, lookups = 0.0, 0
    t0 = time.perf_counter()
    for i in range(a.warmup, a.warmup + a.steps):
        km, lk = step(i)
        kernel_ms += km
        lookups += lk
    elapsed = time.perf_counter() - t0
    avg_ms = kernel_ms / a.steps
    achieved = (lookups / a.steps) * BYTES_PER_LOOKUP / (avg_ms * 1e-3) / 1e9
    log("tracer: %.2f ms/step, kernel %.2f ms, %.1f GB/s algorithmic" % (1000 * elapsed / a.steps, avg_ms, achieved))
    ceiling = random_read_ceiling(ia)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": hbm_traffic("trace_kernel", TRACER_WORKLOAD % a.spp_per_step),
            "kernel": "trace_kernel",
            "kernel_ms": avg_ms, "lookups_per_launch": lookups / a.steps,
            "bytes_per_lookup": BYTES_PER_LOOKUP}
    return {
        "value": a.steps * a.spp_per_step / elapsed, "ms_per_step": 1000.0 * elapsed / a.steps, "dtype": "f32",
        "config": {"workload": "plain delta-tracking tracer, 512^3 Perlin-fBm cloud, %dx%d, 1 spp/frame, "
                               "%d frames/launch, %d bounces" % (sc.W, sc.H, a.spp_per_step, sc.bounces),
                   "width": sc.W, "height": sc.H, "grid": sc.res, "spp_per_step": a.spp_per_step},
        "roofline": with_ceiling(roof, ceiling, lookups / (kernel_ms * 1e-3)),
        "random_read_ceiling": ceiling,
        "extra": {"glookups_per_s": lookups / (kernel_ms * 1e-3) / 1e9, "max_path_steps": sc.st.max_path_steps},
    }


def hbm_traffic(kernel, workload, field="bytes_per_launch"):
    """HBM bytes per launch of `kernel` on `workload`, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/hbm_traffic.json
    (each counter scaled by its reading for a copy kernel of known size)."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    rec = table.get(kernel, {}).get(workload)
    return None if rec is None else rec.get(field)


def time_mlp(a, ia, sc):
    """K timed 1-spp MRPNN frames folded into the running mean in HBM.
    Returns (elapsed s, dict of per-frame averages: HIP-event time of each
    kernel, lookups of the first-vertex and shadow kernels, descriptors)."""
    sc.target.zero()
    for i in range(a.warmup):
        sc.render(i, 1, MLP_SEED, rt=ia.MRPNN, precision=a.mlp_precision, fresh=False)
    keys = ("fv_ms", "shadow_ms", "compact_ms", "gather_ms", "mlp_ms", "descriptors", "fv_lookups", "shadow_lookups")
    acc = dict.fromkeys(keys, 0.0)
    t0 = time.perf_counter()
    for i in range(a.warmup, a.warmup + a.steps):
        sc.render(i, 1, MLP_SEED, rt=ia.MRPNN, precision=a.mlp_precision, fresh=False)
        st = sc.st
        acc["fv_ms"] += st.first_vertex_ms
        acc["shadow_ms"] += st.shadow_ms
        acc["compact_ms"] += st.compact_ms
        acc["gather_ms"] += st.gather_ms
        acc["mlp_ms"] += st.kernel_ms
        acc["descriptors"] += st.scatters
        acc["fv_lookups"] += st.lookups - st.shadow_lookups
        acc["shadow_lookups"] += st.shadow_lookups
    elapsed = time.perf_counter() - t0
    k = {key: v / a.steps for key, v in acc.items()}
    log("MLP renderer: %.2f ms/frame (first vertex %.2f + shadow rays %.2f, compaction %.2f, gather %.2f, network "
        "%.2f ms; %d descriptors = %.1f%% of pixels)"
        % (1000 * elapsed / a.steps, k["fv_ms"], k["shadow_ms"], k["compact_ms"], k["gather_ms"], k["mlp_ms"],
           k["descriptors"], 100.0 * k["descriptors"] / (sc.W * sc.H)))
    return elapsed, k


def mlp_rooflines(k, precision=1, ceiling=None, cells16=False):
    """Rooflines of the kernels of an MRPNN frame; `roofline` is the one that
    takes longest.  Network: FLOPs of the descriptors evaluated (first
    vertices that scatter; escaping pixels are sky and skip gather and network)
    against the dense BF16 (precision 1) or FP8 (precision 2) MFMA peak.
    First vertex and shadow rays: 32 B per trilinear lookup, also given as a
    fraction of the measured random-read ceiling.  Gather: measured HBM bytes
    per launch over its time; its algorithmic 18,400 B per descriptor (12,256 B
    with FP16 cells) come mostly from the caches and are reported beside."""
    nd = k["descriptors"]
    mlp = MLP_FLOP_PER_PASS * nd / (k["mlp_ms"] * 1e-3) / 1e12
    kname, peak = ("mlp_fp8_kernel", FP8_PEAK_TFLOPS) if precision == 2 else ("mlp_bf16_kernel", BF16_PEAK_TFLOPS)
    r = {kname: {"bound": "mfma", "achieved": mlp, "peak": peak, "unit": "TFLOP/s", "frac": mlp / peak,
                 "traffic": hbm_traffic(kname, MLP_WORKLOAD), "kernel": kname, "kernel_ms": k["mlp_ms"],
                 "flop_per_launch": MLP_FLOP_PER_PASS * nd, "descriptors_per_launch": nd}}
    for name, lk, ms in (("first_vertex_kernel", k["fv_lookups"], k["fv_ms"]),
                         ("shadow_kernel", k["shadow_lookups"], k["shadow_ms"])):
        gbs = lk * BYTES_PER_LOOKUP / (ms * 1e-3) / 1e9
        r[name] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                   "traffic": hbm_traffic(name, MLP_WORKLOAD), "kernel": name, "kernel_ms": ms,
                   "lookups_per_launch": lk, "bytes_per_lookup": BYTES_PER_LOOKUP}
        if ceiling:
            r[name] = with_ceiling(r[name], ceiling, lk / (ms * 1e-3))
    # Gather: most of the cell bytes its lookups read are served by L2 and the
    # last-level cache, so its HBM fraction is the measured HBM traffic over its
    # time; the algorithmic rate (cell bytes the lookups read) is given beside it.
    # (the MFMA networks read the wave-tiled gather; the FP32 network the row gather)
    gname = "gather_kernel" if precision == 0 else "gather_tiled_kernel"
    per_desc = GATHER_BYTES_PER_POINT_16 if cells16 else GATHER_BYTES_PER_POINT
    ga_alg = per_desc * nd / (k["gather_ms"] * 1e-3) / 1e9
    traffic = hbm_traffic(gname, MLP_WORKLOAD + ("_cells16" if cells16 else ""))
    ga_hbm = traffic / (k["gather_ms"] * 1e-3) / 1e9 if traffic else None
    r[gname] = {"bound": "hbm", "achieved": ga_hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": ga_hbm / HBM_PEAK_GBS if ga_hbm else None, "traffic": traffic,
                          "kernel": gname, "kernel_ms": k["gather_ms"],
                          "cells": "fp16" if cells16 else "fp32",
                          "achieved_is": "measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                         "scaled by a copy kernel) over the live kernel time",
                          "algorithmic": {"bytes_per_descriptor": per_desc, "descriptors_per_launch": nd,
                                          "gbs": ga_alg, "frac_of_hbm_peak": ga_alg / HBM_PEAK_GBS,
                                          "served_from": "mostly L2 and the last-level cache, not HBM"}}
    dominant = max(r, key=lambda n: r[n]["kernel_ms"])
    return r[dominant], r


def headline_config(sc, a):
    net = "FP8-E4M3" if a.mlp_precision == 2 else ("FP32" if a.mlp_precision == 0 else "BF16")
    cells = "FP16" if a.gather_cells == 16 and a.mlp_precision != 0 else "FP32"
    main = a.network if (a.network != "native" or __import__("internship_amd").has_native_weights()) else "shipped"
    return {"workload": "MRPNN renderer (" + net + " MFMA network, " + main + " weights, descriptor gather on " + cells
                        + " cells), "
                        "512^3 Perlin-fBm cloud seed 1, %dx%d, 1 spp/frame; "
                        "vs plain delta tracking at equal RMSE against a %d-spp FP64 ground-truth image"
                        % (sc.W, sc.H, a.truth_spp),
            "width": sc.W, "height": sc.H, "grid": sc.res, "alpha": 256, "albedo": 0.999, "g": 0.857,
            "bounces": sc.bounces, "truth_spp": a.truth_spp, "gather_cells": cells.lower(), "network": main}


def run_mlp(a, ia, sc):
    """Only the timed MLP-renderer frames (no ground truth, no equal-RMSE search)."""
    main, nets = headline_networks(a, ia)
    net = ia.Network(nets[main])
    sc.vol.set_network(net)
    sc.vol.set_gather_cells(a.gather_cells)
    elapsed, k = time_mlp(a, ia, sc)
    net.close()
    cfg = headline_config(sc, a)
    cfg["workload"] = cfg["workload"].split(";")[0]
    ceiling = random_read_ceiling(ia)
    roof, roofs = mlp_rooflines(k, a.mlp_precision, ceiling, cells16=a.gather_cells == 16)
    return {"value": a.steps / elapsed, "ms_per_step": 1000.0 * elapsed / a.steps,
            "dtype": {0: "f32", 1: "bf16", 2: "fp8-e4m3"}[a.mlp_precision],
            "config": cfg, "roofline": roof, "rooflines": roofs, "random_read_ceiling": ceiling, "kernels": k}


def run_frame(a, ia, sc):
    """Progressive frames (`VolumeRender::Render(target, histo, target2, ...)`):
    the path tracer with the checkerboard on (half the pixels per call), the
    per-pixel histograms, the 11x11 histogram-distance denoiser, ACES and the
    BGRA8 pack, K timed calls."""
    W, H = sc.W, sc.H
    target = ia.DeviceBuffer(W * H * 12)
    histo = ia.DeviceBuffer(W * H * 132)
    bgra = ia.DeviceBuffer(W * H * 4)
    for b in (target, histo, bgra):
        b.zero()
    sc.vol.set_checkerboard(True)
    sc.vol.set_scatter_rate(0.999)
    sc.vol.set_env(0.1)
    ori = 2.5 * np.array([0.75, 0.75, 0.75]) / np.linalg.norm([0.75, 0.75, 0.75])
    right, up = ia.camera_basis(ori, 40.0, W / H)
    light = np.array([0.0, 1.0, 0.3]) / np.linalg.norm([0.0, 1.0, 0.3])
    st = ia.TraceStats()

    def frame(i):
        sc.vol.render_frame(target, histo, bgra, (W, H), ori, up, right, light, alpha=256.0,
                            multi_scatter=sc.bounces, g=0.857, randseed=i, rt=ia.PT, tone=2, denoise=True, stats=st)
        return st.kernel_ms, st.reduce_ms, st.lookups

    for i in range(a.warmup):
        frame(i)
    tr_ms = post_ms = 0.0
    lookups = 0
    t0 = time.perf_counter()
    for i in range(a.warmup, a.warmup + a.steps):
        k, r, lk = frame(i)
        tr_ms += k
        post_ms += r
        lookups += lk
    elapsed = time.perf_counter() - t0
    tr_ms, post_ms = tr_ms / a.steps, post_ms / a.steps
    achieved = (lookups / a.steps) * BYTES_PER_LOOKUP / (tr_ms * 1e-3) / 1e9
    post_gbs = W * H * FRAME_POST_BYTES_PER_PIXEL / (post_ms * 1e-3) / 1e9
    log("progressive frames: %.2f ms/frame (trace %.2f, histogram + denoise + tone + pack %.3f ms)"
        % (1000 * elapsed / a.steps, tr_ms, post_ms))
    return {
        "metric": "1080p progressive path-traced frames/s (checkerboard, histograms, denoise, ACES, BGRA8)",
        "value": a.steps / elapsed, "ms_per_step": 1000.0 * elapsed / a.steps, "dtype": "f32",
        "config": {"workload": "progressive path-traced frames, 512^3 Perlin-fBm cloud, %dx%d, checkerboard "
                               "(half the pixels per frame), histogram denoiser, ACES, BGRA8" % (W, H),
                   "width": W, "height": H, "grid": sc.res, "bounces": sc.bounces},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "trace_kernel",
                     "kernel_ms": tr_ms, "lookups_per_launch": lookups / a.steps,
                     "bytes_per_lookup": BYTES_PER_LOOKUP},
        "kernels": {"trace_ms": tr_ms, "post_ms": post_ms, "post_gbs_compulsory": post_gbs},
    }


def run_train(a, ia, sc):
    """Config 4, the learned part: training data generated on the GPU from
    Perlin-fBm clouds (seeds 100.., thresholds U[0.3,0.5]; per view extinction
    U[64,512], sun on the upper hemisphere, g U[0.7,0.9]; albedo 0.999), points
    sampled as `MeanFreePathSample` does, `GetSample` targets; 5 % of the
    clouds held out.  The network is trained from seeded random weights with
    Adam (lr, batch, epochs) on the BF16 matrix-core trainer; the held-out loss
    and relative L2 of the trained and the shipped weights are compared, and
    (--eval-headline) both networks render the headline scene against its
    ground truth, with plain delta tracking at equal RMSE beside them."""
    from internship_amd import datagen, learn
    views = datagen.config4_views(a.train_clouds, 100, a.views_per_cloud, seed=4)
    ds, gst = datagen.make_config4(views, res=a.train_res, samples_per_view=a.samples_per_view, spp=a.target_spp,
                                   bounces=a.target_bounces, albedo=0.999, seed=1, log=log)
    n_hold = max(1, int(round(0.05 * a.train_clouds)))
    n_train = int(np.argmax(ds.cloud >= a.train_clouds - n_hold))
    log("data: %d samples (%d training, %d held out from %d clouds) in %.1f s: targets %.1f s, %.3g paths/s, "
        "%.3g lookups/s" % (ds.n, n_train, ds.n - n_train, n_hold, gst["total_s"], gst["target_s"],
                            gst["paths"] / max(gst["target_s"], 1e-9), gst["target_lookups"] / max(gst["target_s"], 1e-9)))
    srp = float(ds.aux_host[0, 2])
    mean_y = float(ds.aux_host[:n_train, 3].mean())
    w0 = learn.init_weights(a.init_seed, mean_y, srp)
    B = a.batch
    tr0 = ia.Trainer(w0, max_batch=B, precision=a.train_precision)
    loss0, rel0, _ = learn.evaluate(tr0, ds, n_train, ds.n, B)
    tr0.close()
    shipped = ia.Trainer(ia.load_reference_weights(), max_batch=B, precision=a.train_precision)
    loss_s, rel_s, _ = learn.evaluate(shipped, ds, n_train, ds.n, B)
    shipped.close()
    log("held out before training: random init loss %.5g rel L2 %.4g; shipped weights loss %.5g rel L2 %.4g"
        % (loss0, rel0, loss_s, rel_s))
    w1, hist, t_train = learn.train(ds, n_train, w0, epochs=a.epochs, batch=B, lr=a.lr, seed=a.init_seed,
                                    precision=a.train_precision, log=log, eval_range=(n_train, ds.n),
                                    eval_every=max(1, a.epochs // 10))
    steps = a.epochs * (n_train // B)
    sps = steps * B / t_train
    if a.save_weights:
        np.save(a.save_weights, w1)
    res = {
        "metric": "MRPNN training from seeded random weights: samples/s (forward + backward + Adam, GPU) and "
                  "held-out relative L2", "unit": "samples/s",
        "value": sps, "ms_per_step": 1000.0 * t_train / steps, "dtype": "bf16" if a.train_precision else "f32",
        "config": {"workload": "config 4: %d Perlin-fBm clouds (seeds 100-%d, %d^3) x %d views, %d samples per view, "
                               "GetSample targets %d spp / %d bounces, albedo 0.999; %d clouds held out; Adam lr %g, "
                               "batch %d, %d epochs from seeded random weights (seed %d)"
                               % (a.train_clouds, 99 + a.train_clouds, a.train_res, a.views_per_cloud,
                                  a.samples_per_view, a.target_spp, a.target_bounces, n_hold, a.lr, B, a.epochs,
                                  a.init_seed),
                   "samples": ds.n, "train_samples": n_train, "heldout_samples": ds.n - n_train,
                   "batch": B, "lr": a.lr, "epochs": a.epochs, "precision": a.train_precision},
        "datagen": {k: (float(v) if isinstance(v, (int, float, np.floating)) else v) for k, v in gst.items()},
        "heldout": {"random_init": {"loss": loss0, "rel_l2": rel0}, "shipped_weights": {"loss": loss_s, "rel_l2": rel_s},
                    "trained": {"loss": hist[-1][2], "rel_l2": hist[-1][3]}},
        "history": [{"epoch": h[0], "train_loss": h[1], "heldout_loss": h[2], "heldout_rel_l2": h[3]} for h in hist],
        "roofline": {"bound": "mfma", "achieved": sps * 3 * MLP_FLOP_PER_PASS / 1e12, "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": sps * 3 * MLP_FLOP_PER_PASS / 1e12 / BF16_PEAK_TFLOPS,
                     "traffic": None, "kernel": "train_fwd_mfma + train_bwd_mfma + wgrad + adam + gather_rows",
                     "flop_per_sample": 3 * MLP_FLOP_PER_PASS},
    }
    if a.eval_headline:
        res["headline_scene"] = eval_networks_on_headline(a, ia, {"shipped": ia.load_reference_weights(),
                                                                  "trained": w1})
    return res


def eval_networks_on_headline(a, ia, nets):
    """RMSE of 1-spp BF16 MRPNN frames of the headline scene with each network
    against the 4096-spp ground truth, and the delta-tracking spp at equal RMSE."""
    sc = Scene(ia, 512, 1920, 1080, 512)
    ref = ground_truth_image(ia, sc, a.truth_spp)
    out = {}
    for name, w in nets.items():
        net = ia.Network(w)
        sc.vol.set_network(net)
        el, k = time_mlp(a, ia, sc)
        sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=1)
        r = rmse(sc.image(), ref)
        net.close()
        n_eq, hist, reached = equal_rmse_search(sc, r, ref, a.dt_max_spp)
        t_dt, _, _ = dt_time(sc, n_eq)
        out[name] = {"rmse": r, "frames_per_s": a.steps / el, "dt_spp_at_equal_rmse": n_eq, "dt_reached": reached,
                     "speedup_vs_dt": (a.steps / el) * t_dt}
        log("headline scene, %s network: RMSE %.5g, DT at equal RMSE %d spp, speed-up %.1fx"
            % (name, r, n_eq, out[name]["speedup_vs_dt"]))
    return out


def run_serve(a, ia, sc):
    """Batched serving: 16 Perlin-fBm clouds with resolutions drawn from
    {128,192,256,384,512}^3 (seeds 200-215) x 4 orbit cameras at 1280x720 =
    64 views per batch, each call (mrp_serve_views) rendering all of them:
    BF16 MRPNN (the native network), a delta-tracking fallback for pixels whose
    sample variance exceeds --serve-threshold (one persistent launch across
    the scenes), histogram denoise and ACES pack.  A step is one batch; its
    latency is the per-frame latency of every view in it.  RMSE per scene is
    over linear RGB of camera 0 against a --serve-truth-spp delta-tracking
    image with FP64 sums (the GPU tracer, which is bit-exact with the oracle)."""
    from internship_amd.serving import serving_scenes, view_params
    W, H = a.serve_width, a.serve_height
    scenes = serving_scenes(16, 4)
    which, nets = headline_networks(a, ia)
    net = ia.Network(nets[which])
    vols, views, vlist = [], [], []
    t0 = time.perf_counter()
    for k, s in enumerate(scenes):
        grid = ia.DeviceBuffer(s.res ** 3 * 4)
        ia.generate_cloud(s.res, seed=s.seed, threshold=s.threshold, device_out=grid)
        vol = ia.VolumeRender(s.res)
        vol.set_datas(grid)
        vol.update()
        grid.free()
        vol.set_network(net)
        vols.append(vol)
        for cam in s.cameras:
            views.append(view_params(s, cam, W, H, g=0.857, albedo=0.99, multi=a.bounces, seed=k, precision=1))
            vlist.append(vol)
    log("%d scenes (%s) prepared in %.1f s" % (len(scenes), ",".join(str(s.res) for s in scenes),
                                               time.perf_counter() - t0))
    out = ia.ServeBatch(len(views), W, H)
    kw = dict(mrpnn_spp=a.serve_spp, var_threshold=a.serve_threshold, fallback_spp=a.serve_fallback_spp,
              denoise=True, tone=2)
    for _ in range(max(1, a.warmup)):
        st = out.run(vlist, views, **kw)
    lat, stats = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        st = out.run(vlist, views, **kw)
        lat.append(time.perf_counter() - t)
        stats.append(st.asdict())
    elapsed = time.perf_counter() - t0
    lat_ms = np.array(lat) * 1000.0
    n_views = len(views)
    mean = {k: float(np.mean([d[k] for d in stats])) for k in stats[0]}
    log("serving: %d batches of %d views in %.2f s: %.1f frames/s, batch latency p50 %.1f ms, p99 %.1f ms; "
        "device %.1f ms (MRPNN %.1f, accumulate %.2f, fallback %.1f, denoise %.1f); fallback pixels %.0f (%.2f%%)"
        % (a.steps, n_views, elapsed, n_views * a.steps / elapsed, np.percentile(lat_ms, 50),
           np.percentile(lat_ms, 99), mean["total_ms"], mean["mrpnn_ms"], mean["accumulate_ms"], mean["fallback_ms"],
           mean["denoise_ms"], mean["fallback_pixels"], 100.0 * mean["fallback_pixels"] / (n_views * W * H)))
    rgb, bgra, flags = out.download()
    # MRPNN alone (no fallback, no denoise) for comparison, same views
    out.run(vlist, views, mrpnn_spp=a.serve_spp, var_threshold=1e30, fallback_spp=1, denoise=False, tone=2,
            flags=False)
    rgb_nn = out.rgb.download(np.float32, (n_views, H, W, 3))
    bgra_nn = out.bgra.download(np.uint32, (n_views, H, W))
    per_scene = []
    if a.serve_truth_spp > 0:
        t0 = time.perf_counter()
        for k, s in enumerate(scenes):
            i = 4 * k
            p = ia.RenderParams.from_buffer_copy(views[i])
            p.render_type, p.spp_begin, p.spp_count = ia.PT, 0, a.serve_truth_spp
            tgt = ia.DeviceBuffer(W * H * 12)
            tgt.zero()
            s64 = ia.DeviceBuffer(W * H * 24)
            s64.zero()
            vlist[i].render_dev(p, tgt, s64)
            ref = s64.download(np.float64, (H, W, 3)) / a.serve_truth_spp
            tgt.free()
            s64.free()
            d = {"scene": k, "res": s.res, "seed": s.seed, "alpha": round(s.alpha, 2),
                 "rmse_served": rmse(rgb[i], ref), "rmse_mrpnn_only": rmse(rgb_nn[i], ref),
                 "luma_rmse100_served_denoised": luma_rmse100_packed(bgra[i], ref),
                 "luma_rmse100_mrpnn_only": luma_rmse100_packed(bgra_nn[i], ref),
                 "fallback_frac": float(flags[i].mean())}
            per_scene.append(d)
            log("  scene %2d (%3d^3, alpha %6.1f): RMSE served %.5f, MRPNN only %.5f; luma RMSEx100 served+denoise "
                "%.3f, MRPNN only %.3f; fallback %.2f%% (%.1f s)"
                % (d["scene"], d["res"], d["alpha"], d["rmse_served"], d["rmse_mrpnn_only"],
                   d["luma_rmse100_served_denoised"], d["luma_rmse100_mrpnn_only"], 100 * d["fallback_frac"],
                   time.perf_counter() - t0))
    net.close()
    for v in vols:
        v.close()
    return {
        "metric": "720p served frames/s over 16 scenes x 4 cameras (batched: MRPNN + DT fallback + denoise)",
        "unit": "frames/s", "value": n_views * a.steps / elapsed, "ms_per_step": 1000.0 * elapsed / a.steps,
        "dtype": "bf16",
        "config": {"workload": "16 Perlin-fBm clouds (128..512^3, seeds 200-215) x 4 orbit cameras, %dx%d, %d BF16 "
                               "MRPNN spp, delta-tracking fallback (%d spp, %d bounces) above variance %g, denoise, "
                               "%s network; one batch of %d views per step" % (W, H, a.serve_spp,
                                                                              a.serve_fallback_spp, a.bounces,
                                                                              a.serve_threshold, which, n_views),
                   "scenes": len(scenes), "cameras": 4, "width": W, "height": H, "network": which},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p90": float(np.percentile(lat_ms, 90)),
                       "p99": float(np.percentile(lat_ms, 99)), "max": float(lat_ms.max()),
                       "note": "per batch = per frame of every view in it"},
        "device_ms": mean, "per_scene": per_scene, "truth_spp": a.serve_truth_spp,
    }


def headline_networks(a, ia):
    """{name: weights} of the networks the headline renders with, and the one
    it times: the network this library trained from seeded random weights on
    its config-4 data (native, when committed) and the shipped NNWeight.cuh
    weights."""
    nets = {"shipped": ia.load_reference_weights()}
    if ia.has_native_weights():
        nets["native"] = ia.load_native_weights()
    return (a.network if a.network in nets else "shipped"), nets


def run_headline(a, ia, sc):
    main, nets = headline_networks(a, ia)
    net = ia.Network(nets[main])
    sc.vol.set_network(net)
    sc.vol.set_gather_cells(a.gather_cells)
    P = sc.W * sc.H
    ref = ground_truth_image(ia, sc, a.truth_spp)  # untimed
    elapsed, k = time_mlp(a, ia, sc)
    fps_mlp = a.steps / elapsed
    # quality of one 1-spp frame with the BF16, FP32 and FP8 networks
    sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=1)
    img_mlp = sc.image()
    rmse_mlp = rmse(img_mlp, ref)
    # each network's time from a second call (the first launch of a kernel loads its code)
    for _ in range(2):
        sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=0)
    fp32_ms = sc.st.kernel_ms
    rmse_mlp32 = rmse(sc.image(), ref)
    for _ in range(2):
        sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=2)
    fp8_ms = sc.st.kernel_ms
    rmse_mlp8 = rmse(sc.image(), ref)
    log("MLP 1-spp RMSE: BF16 %.5g, FP32 %.5g, FP8 %.5g (networks: FP32 %.2f ms, FP8 %.2f ms)"
        % (rmse_mlp, rmse_mlp32, rmse_mlp8, fp32_ms, fp8_ms))
    # the other cell format of the BF16 gather, timed the same way (a named variant, not the headline)
    other = 16 if a.gather_cells == 32 else 32
    sc.vol.set_gather_cells(other)
    el_o, k_o = time_mlp(a, ia, sc)
    sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=1)
    variant = {"gather_cells": "fp%d" % other, "frames_per_s": a.steps / el_o, "ms_per_frame": 1000.0 * el_o / a.steps,
               "gather_ms": k_o["gather_ms"], "rmse_mlp_bf16": rmse(sc.image(), ref)}
    sc.vol.set_gather_cells(a.gather_cells)
    log("BF16 renderer with FP%d gather cells: %.1f frames/s (gather %.3f ms), RMSE %.5g"
        % (other, variant["frames_per_s"], variant["gather_ms"], variant["rmse_mlp_bf16"]))
    # the plain tracer at equal RMSE
    n_eq, hist, reached = equal_rmse_search(sc, rmse_mlp, ref, a.dt_max_spp)
    t_dt, dt_kernel_ms, dt_lookups = dt_time(sc, n_eq)
    fps_dt = 1.0 / t_dt
    sc.render(0, 1, DT_SEED)
    img_dt1 = sc.image()
    log("DT %d spp: %.2f ms/frame -> speed-up %.2fx" % (n_eq, 1000 * t_dt, fps_mlp / fps_dt))
    dt_achieved = dt_lookups * BYTES_PER_LOOKUP / (dt_kernel_ms * 1e-3) / 1e9
    net.close()
    # the other network(s) at equal RMSE (same kernels, so the same frames/s)
    others = {}
    for name, w in nets.items():
        if name == main:
            continue
        onet = ia.Network(w)
        sc.vol.set_network(onet)
        sc.render(0, 1, MLP_SEED, rt=ia.MRPNN, precision=1)
        r_o = rmse(sc.image(), ref)
        onet.close()
        n_o, hist_o, reached_o = equal_rmse_search(sc, r_o, ref, a.dt_max_spp)
        t_o, _, _ = dt_time(sc, n_o)
        others[name] = {"rmse_mlp_bf16": r_o, "dt_spp": n_o, "dt_reached_target": reached_o,
                        "rmse_dt": hist_o.get(n_o), "fps_dt_at_equal_rmse": 1.0 / t_o, "speedup": fps_mlp * t_o}
        log("%s network: RMSE %.5g, DT at equal RMSE %d spp, speed-up %.1fx" % (name, r_o, n_o, fps_mlp * t_o))
    ceiling = random_read_ceiling(ia)
    roof, roofs = mlp_rooflines(k, a.mlp_precision, ceiling, cells16=a.gather_cells == 16)
    # the two rooflines the targets name: the tracer against HBM, the network against MFMA
    tr_traffic = hbm_traffic("trace_kernel", TRACER_WORKLOAD % 16)
    tr_lookups16 = hbm_traffic("trace_kernel", TRACER_WORKLOAD % 16, "lookups_per_launch")
    bpl = tr_traffic / tr_lookups16 if tr_traffic and tr_lookups16 else None
    tracer_hbm = dt_lookups * bpl / (dt_kernel_ms * 1e-3) / 1e9 if bpl else None
    mlp_key = "mlp_fp8_kernel" if a.mlp_precision == 2 else "mlp_bf16_kernel"
    north_star = {
        "tracer_hbm": {"kernel": "trace_kernel", "launch": "%d spp at 1920x1080 (equal RMSE)" % n_eq,
                       "kernel_ms": dt_kernel_ms, "lookups_per_launch": dt_lookups,
                       "algorithmic_gbs": dt_achieved, "algorithmic_frac": dt_achieved / HBM_PEAK_GBS,
                       "hbm_bytes_per_lookup": bpl, "hbm_gbs": tracer_hbm,
                       "hbm_frac": tracer_hbm / HBM_PEAK_GBS if tracer_hbm else None,
                       "frac_of_random_read_ceiling": dt_lookups / (dt_kernel_ms * 1e-3) / ceiling["reads_per_s"],
                       "target_frac": 0.5,
                       "note": "algorithmic = 32 B per trilinear lookup; hbm = the HBM bytes per lookup measured "
                               "by rocprofv3 on the 16-spp launch (whole 128-byte lines)"},
        "mlp_mfma": dict(roofs[mlp_key], target_frac=0.4) if mlp_key in roofs else None}
    return {
        "value": fps_mlp, "ms_per_step": 1000.0 * elapsed / a.steps,
        "dtype": {0: "f32", 1: "bf16", 2: "fp8-e4m3"}[a.mlp_precision],
        "config": headline_config(sc, a),
        "roofline": roof, "rooflines": roofs, "north_star": north_star, "random_read_ceiling": ceiling,
        "equal_rmse": {"rmse_mlp_bf16": rmse_mlp, "rmse_mlp_fp32": rmse_mlp32, "rmse_mlp_fp8": rmse_mlp8,
                       "dt_spp": n_eq,
                       "dt_reached_target": reached, "rmse_dt": hist.get(n_eq), "fps_mlp": fps_mlp,
                       "fps_dt_at_equal_rmse": fps_dt, "speedup": fps_mlp / fps_dt,
                       "dt_rmse_by_spp": {str(k): v for k, v in sorted(hist.items())},
                       "luma_rmse100_mlp": luma_rmse100(img_mlp, ref),
                       "luma_rmse100_dt1": luma_rmse100(img_dt1, ref)},
        "variant": variant,
        "network": {"timed": main, "weights": "trained by this library from seeded random weights on its config-4 "
                                               "data (data/mrpnn_weights_native.*)" if main == "native" else
                    "core/NNWeight.cuh (shipped)", "other_networks": others},
        "kernels": dict(k, mlp_fp32_ms=fp32_ms, mlp_fp8_ms=fp8_ms, tracer_kernel_ms_at_equal_rmse=dt_kernel_ms),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", choices=["headline", "mlp", "tracer", "frame", "train", "serve"], default="headline")
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bounces", type=int, default=512)
    ap.add_argument("--truth-spp", type=int, default=4096)
    ap.add_argument("--dt-max-spp", type=int, default=4096)
    ap.add_argument("--spp-per-step", type=int, default=1, help="tracer mode: 1-spp frames per launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--network", choices=["native", "shipped"], default="native",
                    help="headline/mlp modes: the network trained by this library (native, when its weights are "
                         "committed) or the shipped NNWeight.cuh weights")
    ap.add_argument("--gather-cells", type=int, choices=[32, 16], default=32,
                    help="headline/mlp modes: cell format of the BF16/FP8 gather (32 FP32, 16 FP16)")
    ap.add_argument("--mlp-precision", type=int, choices=[0, 1, 2], default=1,
                    help="headline/mlp modes: network of the timed frames (0 FP32, 1 BF16, 2 FP8-E4M3)")
    ap.add_argument("--batch", type=int, default=16384, help="train mode: samples per step")
    ap.add_argument("--serve-width", type=int, default=1280)
    ap.add_argument("--serve-height", type=int, default=720)
    ap.add_argument("--serve-spp", type=int, default=2, help="serve mode: MRPNN samples per pixel")
    ap.add_argument("--serve-threshold", type=float, default=0.05, help="serve mode: fallback variance threshold")
    ap.add_argument("--serve-fallback-spp", type=int, default=32, help="serve mode: delta-tracking spp of a fallback")
    ap.add_argument("--serve-truth-spp", type=int, default=4096, help="serve mode: ground-truth spp (0: none)")
    ap.add_argument("--lr", type=float, default=1e-3, help="train mode: Adam learning rate")
    ap.add_argument("--epochs", type=int, default=20, help="train mode: epochs over the training samples")
    ap.add_argument("--init-seed", type=int, default=0, help="train mode: seed of the random initial weights")
    ap.add_argument("--views-per-cloud", type=int, default=4, help="train mode: light/extinction/g draws per cloud")
    ap.add_argument("--samples-per-view", type=int, default=2048)
    ap.add_argument("--train-res", type=int, default=256, help="train mode: cloud resolution")
    ap.add_argument("--save-weights", default="", help="train mode: .npy path for the trained weights")
    ap.add_argument("--eval-headline", action="store_true",
                    help="train mode: render the headline scene with the shipped and the trained network")
    ap.add_argument("--train-precision", type=int, choices=[0, 1], default=1,
                    help="train mode: 0 FP32 forward/backward, 1 BF16 matrix cores (FP32 master weights)")
    ap.add_argument("--train-clouds", type=int, default=64)
    ap.add_argument("--target-spp", type=int, default=256)
    ap.add_argument("--target-bounces", type=int, default=512)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (reference kernel for HBM counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(a frame would shard by image tiles, training by samples)", file=sys.stderr)
        return 2

    import internship_amd as ia
    lib = ia.load()
    if a.copy_calibration_mib > 0:
        nb = a.copy_calibration_mib << 20
        src, dst = ia.DeviceBuffer(nb), ia.DeviceBuffer(nb)
        src.zero()
        ia.check(lib.mrp_memcpy_d2d(dst.ptr, src.ptr, ctypes.c_size_t(nb)))
        log("calibration copy of %d MiB done" % a.copy_calibration_mib)
        src.free()
        dst.free()
    sc = Scene(ia, a.res, a.width, a.height, a.bounces) if a.mode not in ("train", "serve") else None
    res = {"headline": run_headline, "mlp": run_mlp, "tracer": run_tracer, "frame": run_frame,
           "train": run_train, "serve": run_serve}[a.mode](a, ia, sc)
    base = None
    if not a.no_cpu_baseline and a.mode not in ("train", "serve"):
        log("timing the CPU oracle on a bounded sample")
        base = cpu_baseline(sc)
        log("cpu baseline %.4g frames/s (%s)" % (base["value"], base["sample"]))
    out = {"metric": res.pop("metric", METRIC), "value": res.pop("value"), "unit": res.pop("unit", "frames/s"),
           "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": res.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": res.pop("dtype"), "data": "synthetic"}
    out.update(res)
    out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
