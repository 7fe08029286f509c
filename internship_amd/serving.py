"""Batched serving of several scenes (mrp_serve_views): the scene set of the
serving workload and its cameras.

The workload: 16 procedural clouds (Perlin fBm, seeds 200-215) with
resolutions drawn from {128, 192, 256, 384, 512}^3, each seen from 4 orbit
cameras at 1280x720, rendered in one batch with the trained BF16 network, a
delta-tracking fallback for pixels whose sample variance exceeds a threshold,
and the histogram denoise.  The draws come from numpy's PCG64 with seed 5 so
that the scene set is fixed.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import RenderParams, Vec3, camera_basis

RESOLUTIONS = (128, 192, 256, 384, 512)
SCENE_SEEDS = tuple(range(200, 216))


@dataclass
class Scene:
    seed: int
    res: int
    threshold: float  # the cloud's density threshold
    alpha: float      # extinction scale
    light: np.ndarray
    cameras: list     # (ori, up, right) per view


def orbit_camera(azimuth_deg, elevation_deg, distance=2.5, fov_deg=40.0, aspect=16.0 / 9.0):
    """(ori, up, right) of a camera at `distance` from the volume centre
    looking at it, with the image plane scaled for a vertical fov."""
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    ori = distance * np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    right, up = camera_basis(ori, fov_deg, aspect)
    return ori.astype(np.float32), up, right


def serving_scenes(n_scenes=16, views_per_scene=4, rng_seed=5):
    rng = np.random.default_rng(rng_seed)
    scenes = []
    for k in range(n_scenes):
        res = int(rng.choice(RESOLUTIONS))
        thr = float(rng.uniform(0.3, 0.5))
        alpha = float(rng.uniform(64.0, 256.0))
        th = rng.uniform(0, 2 * np.pi)
        ph = rng.uniform(0.35, 1.35)  # sun above the horizon
        light = np.array([np.cos(ph) * np.cos(th), np.sin(ph), np.cos(ph) * np.sin(th)])
        cams = [orbit_camera(rng.uniform(0, 360), rng.uniform(-20, 50)) for _ in range(views_per_scene)]
        scenes.append(Scene(SCENE_SEEDS[k % len(SCENE_SEEDS)], res, thr, alpha, light, cams))
    return scenes


def view_params(scene, camera, width=1280, height=720, g=0.857, albedo=0.99, multi=512, seed=0, precision=1,
                sky=(0.1, 0.1, 0.1), light_color=(1.0, 1.0, 1.0)):
    """RenderParams of one view (render type MRPNN)."""
    ori, up, right = camera
    return RenderParams(int(width), int(height), Vec3.of(ori), Vec3.of(up), Vec3.of(right), Vec3.of(scene.light),
                        Vec3.of(light_color), Vec3.of(sky), Vec3.of((albedo, albedo, albedo)), float(g),
                        float(scene.alpha), 1.0, int(multi), 0, 1, int(seed), 2, int(precision))
