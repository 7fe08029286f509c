"""Host tone mapping and BMP output of `Camera::RenderToFile`.

Follows `Camera::Gamma` / `Camera::ACES` (core/camera.cu:66-160, the same
filmic curve as the device `ACES`, core/render.cuh:640-732) and `SaveBMP`
(core/camera.cu:10-64: 32-bit BGRA, negative height = top-down rows).
The colour matrices are the published ACES/sRGB ones and must stay equal to
the reference's for its images to match.
"""
import struct

import numpy as np

_SRGB_2_AP0 = np.array([[0.4397010, 0.3829780, 0.1773350], [0.0897923, 0.8134230, 0.0967616],
                        [0.0175440, 0.1115440, 0.8707040]], dtype=np.float32)
_AP0_2_AP1 = np.array([[1.4514393161, -0.2365107469, -0.2149285693], [-0.0765537734, 1.1762296998, -0.0996759264],
                       [0.0083161484, -0.0060324498, 0.9977163014]], dtype=np.float32)
_AP1_2_XYZ = np.array([[0.6624541811, 0.1340042065, 0.1561876870], [0.2722287168, 0.6740817658, 0.0536895174],
                       [-0.0055746495, 0.0040607335, 1.0103391003]], dtype=np.float32)
_XYZ_2_AP1 = np.array([[1.6410233797, -0.3248032942, -0.2364246952], [-0.6636628587, 1.6153315917, 0.0167563477],
                       [0.0117218943, -0.0082844420, 0.9883948585]], dtype=np.float32)
_D60_2_D65 = np.array([[0.98722400, -0.00611327, 0.0159533], [-0.00759836, 1.00186000, 0.0053302],
                       [0.00307257, -0.00509595, 1.0816800]], dtype=np.float32)
_XYZ_2_REC709 = np.array([[3.2409699419, -1.5373831776, -0.4986107603], [-0.9692436363, 1.8759675015, 0.0415550574],
                          [0.0556300797, -0.2039769589, 1.0569715142]], dtype=np.float32)
_AP1_RGB2Y = np.array([0.272229, 0.674082, 0.0536895], dtype=np.float32)


def gamma_np(c):
    return np.power(np.maximum(c, 0.0), 1.0 / 2.2).astype(np.float32)


def _mat(m, x):
    return x @ m.T


def aces_np(c: np.ndarray) -> np.ndarray:
    c = np.asarray(c, dtype=np.float32)
    acescg = _mat(_AP0_2_AP1, _mat(_SRGB_2_AP0, c))
    y = acescg @ _AP1_RGB2Y
    acescg = y[..., None] * (1 - 0.96) + acescg * 0.96
    a, b, cc, d, e = 278.5085, 10.7772, 293.6045, 88.7122, 80.6889
    x = acescg
    post = (x * (x * a + b)) / (x * (x * cc + d) + e)
    xyz = _mat(_AP1_2_XYZ, post)
    div = np.maximum(xyz.sum(-1), 1e-4)
    xyy = np.stack([xyz[..., 0] / div, xyz[..., 1] / div, xyz[..., 1]], -1)
    xyy[..., 2] = np.power(np.clip(xyy[..., 2], 0.0, 65504.0), 0.9811)
    m = xyy[..., 2] / np.maximum(xyy[..., 1], 1e-4)
    xyz = np.stack([xyy[..., 0] * m, xyy[..., 2], (1.0 - xyy[..., 0] - xyy[..., 1]) * m], -1)
    lin = _mat(_XYZ_2_AP1, xyz)
    y = lin @ _AP1_RGB2Y
    lin = y[..., None] * (1 - 0.93) + lin * 0.93
    xyz = _mat(_D60_2_D65, _mat(_AP1_2_XYZ, lin))
    return gamma_np(_mat(_XYZ_2_REC709, xyz)).astype(np.float32)


def tone_map(img, tone="aces"):
    if tone is None or tone == "none":
        return np.asarray(img, dtype=np.float32)
    if tone == "gamma":
        return gamma_np(img)
    return aces_np(img)


def to_bgra8(img: np.ndarray) -> np.ndarray:
    """Saturate and pack to 8-bit BGRA (alpha 255)."""
    v = (np.clip(img, 0.0, 1.0) * 255.0).astype(np.uint8)
    h, w, _ = v.shape
    out = np.empty((h, w, 4), dtype=np.uint8)
    out[..., 0], out[..., 1], out[..., 2], out[..., 3] = v[..., 2], v[..., 1], v[..., 0], 255
    return out


def save_bmp(img: np.ndarray, path: str):
    bgra = to_bgra8(img)
    h, w, _ = bgra.shape
    size = 54 + w * h * 4
    # the 54-byte header of `SaveBMP` (core/camera.cu:10-64): 32-bit, top-down (negative height)
    header = struct.pack("<2sIHHIIiiHHIIiiII", b"BM", size, 0, 0, 54, 40, w, -h, 1, 32, 0, 0x100000, 0, 0, 0, 0)
    with open(path, "wb") as f:
        f.write(header)
        f.write(bgra.tobytes())
