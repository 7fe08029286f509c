"""Training the MRPNN network from seeded random weights on the GPU.

The reference trains a different network in PyTorch (train/train.py:131-254,
Adam lr 1e-3, train/train.py:536-546) and ships the weights of the MRPNN one
(core/NNWeight.cuh).  Here the MRPNN network of the renderer is trained by the
library's own forward, backward and Adam kernels (mrp_trainer_*):

- init_weights: every layer's weights and biases uniform in
  [-1/sqrt(fan_in), 1/sqrt(fan_in)] (the default of the reference trainer's
  `nn.Linear` layers), seeded; the last bias so that the initial prediction is
  the mean target;
- train: epochs over a Dataset in HBM, each a fresh permutation; a batch's
  rows are gathered into contiguous staging buffers on the GPU
  (mrp_gather_rows) and stepped with Adam;
- evaluate: loss and relative L2 of a contiguous range of samples (the
  held-out clouds) with the trainer's forward.
"""
import ctypes
import time

import numpy as np

from . import DeviceBuffer, Trainer, _lib, check, weight_manifest

N_WEIGHTS = 49673

# fan-in of the layers without a bias (squeeze-excitation and final gates):
# SE blocks 8 -> 8 -> 3 (pooled mean/max of 3 streams, g, gamma), the final
# gate 75 -> 16 -> 6 (core/radiancePredict.cu:240-294,436-446)
_NO_BIAS_FAN_IN = {"1W": 8, "2W": 8}
_FIN_FAN_IN = {"LSEFin1W": 75, "LSEFin2W": 16}


def layer_fan_in():
    """{weight array name: fan-in} from the manifest (row-major [out][in])."""
    arrays = weight_manifest()["arrays"]
    by_name = {a["name"]: a for a in arrays}
    out = {}
    for a in arrays:
        n = a["name"]
        if not n.endswith("W"):
            continue
        b = by_name.get(n[:-1] + "B")
        if b is not None:
            out[n] = a["count"] // b["count"]
        elif n in _FIN_FAN_IN:
            out[n] = _FIN_FAN_IN[n]
        else:
            out[n] = _NO_BIAS_FAN_IN[n[-2:]]
    return out


def init_weights(seed, mean_target=None, srp=None):
    """Seeded random weights in the layout of core/NNWeight.cuh."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_WEIGHTS, np.float32)
    fan = layer_fan_in()
    for a in weight_manifest()["arrays"]:
        n, o, c = a["name"], a["offset"], a["count"]
        fi = fan[n] if n.endswith("W") else fan[n[:-1] + "W"]
        bound = 1.0 / np.sqrt(fi)
        w[o:o + c] = rng.uniform(-bound, bound, c).astype(np.float32)
    if mean_target is not None and srp:
        # p = max(exp(P) - 1, 0) * srp with P = relu(W x + b): start at the mean target
        for a in weight_manifest()["arrays"]:
            if a["name"] == "LX12B":
                w[a["offset"]] = np.float32(np.log1p(mean_target / srp))
    return w


def _view(buf, byte_offset):
    from .datagen import DeviceView
    return DeviceView(buf, byte_offset)


def evaluate(trainer, ds, begin, end, batch):
    """(mean loss, relative L2 sum (p-y)^2 / sum y^2, predictions) of samples
    [begin, end) with the trainer's current weights (no update)."""
    n = end - begin
    preds = np.zeros(n, np.float32)
    loss_sum = 0.0
    for b0 in range(0, n, batch):
        bs = min(batch, n - b0)
        f = _view(ds.feat, (begin + b0) * 576 * 4)
        x = _view(ds.aux, (begin + b0) * 16)
        lo = ctypes.c_double()
        check(trainer.lib.mrp_trainer_gradient(trainer.h, bs, f.ptr, x.ptr, None, ctypes.byref(lo)))
        loss_sum += lo.value * bs
        preds[b0:b0 + bs] = trainer.predictions(bs)
    y = ds.aux_host[begin:end, 3].astype(np.float64)
    rel = float(np.sum((preds - y) ** 2) / np.sum(y * y))
    return loss_sum / n, rel, preds


def train(ds, n_train, weights0, epochs=20, batch=16384, lr=1e-3, seed=0, precision=1, log=None, eval_range=None,
          eval_every=0):
    """Adam epochs over samples [0, n_train) of the Dataset.  Returns (weights,
    history: list of (epoch, mean train loss, eval loss, eval rel L2), seconds)."""
    lib = _lib.load()
    tr = Trainer(weights0, max_batch=batch, precision=precision)
    stage_f = DeviceBuffer(batch * 576 * 4)
    stage_a = DeviceBuffer(batch * 16)
    perm_b = DeviceBuffer(max(1, n_train) * 4)
    rng = np.random.default_rng(seed)
    steps = n_train // batch
    if steps == 0:
        raise ValueError("fewer training samples than one batch")
    hist = []
    t_train = 0.0
    try:
        for e in range(epochs):
            perm_b.upload(rng.permutation(n_train).astype(np.uint32))
            t0 = time.perf_counter()
            losses = 0.0
            for b in range(steps):
                idx = _view(perm_b, b * batch * 4)
                check(lib.mrp_gather_rows(stage_f.ptr, ds.feat.ptr, idx.ptr, batch, 576 * 4, n_train))
                check(lib.mrp_gather_rows(stage_a.ptr, ds.aux.ptr, idx.ptr, batch, 16, n_train))
                losses += tr.step(stage_f, stage_a, batch, lr)
            t_train += time.perf_counter() - t0
            row = [e + 1, losses / steps, None, None]
            if eval_range and (eval_every and ((e + 1) % eval_every == 0) or e + 1 == epochs):
                row[2], row[3], _ = evaluate(tr, ds, eval_range[0], eval_range[1], batch)
            hist.append(tuple(row))
            if log:
                log("epoch %d/%d: train loss %.5g%s (%.1f s training)"
                    % (e + 1, epochs, row[1],
                       "" if row[2] is None else ", held-out loss %.5g, rel L2 %.4g" % (row[2], row[3]), t_train))
        w = tr.weights()
    finally:
        tr.close()
    return w, hist, t_train
