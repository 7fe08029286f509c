"""Reduced-precision rounding in numpy (float64 in and out), to emulate the
MFMA networks on the CPU: OCP FP8-E4M3 and BF16, round to nearest even."""
import numpy as np


def round_significand(v, mbits, emin):
    """Round to `mbits` fraction bits, with exponents below `emin` treated as
    subnormal (fixed quantum 2^(emin - mbits))."""
    v = np.asarray(v, np.float64)
    a = np.abs(v)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    q = np.exp2(np.maximum(e, emin) - mbits)
    return np.round(v / q) * q  # numpy rounds halves to even


def e4m3(v):
    """OCP E4M3: 3 fraction bits, exponent bias 7 (normal exponents >= -6), largest finite 448."""
    return np.clip(round_significand(v, 3, -6), -448.0, 448.0)


def bf16(v):
    return round_significand(v, 7, -126)


def pow2_scale(m):
    """2^(7 - floor(log2 m)): m scaled by it lies in [128, 256); 1 where m == 0."""
    m = np.asarray(m, np.float64)
    return np.where(m > 0, np.exp2(7 - np.floor(np.log2(np.where(m > 0, m, 1.0)))), 1.0)


# Layers the kernels evaluate in FP32 on the vector ALUs, not on MFMA.
VALU_LAYERS = ("LGGSW", "LX12W")


def fp8_layer(name, W, x):
    """The FP8 network's rounding of one layer: weights with one power-of-two
    scale per layer, inputs with one per layer and descriptor (row of x)."""
    if name in VALU_LAYERS:
        return W, x
    sw = pow2_scale(np.abs(W).max())
    sx = pow2_scale(np.abs(x).max(axis=1))[:, None]
    return e4m3(W * sw) / sw, e4m3(x * sx) / sx


def bf16_layer(name, W, x):
    """The BF16 network's rounding of one layer (weights and inputs, no scales)."""
    if name in VALU_LAYERS:
        return W, x
    return bf16(W), bf16(x)
