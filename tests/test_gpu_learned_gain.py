"""What the learned part gains over the method without it.

The network is fine-tuned on the GPU by this library's own kernels (data:
procedural clouds, descriptors from the gather kernel, `GetSample` targets
from the GPU path tracer; training: mrp_trainer forward/backward/Adam,
starting from the shipped weights).  Then a held-out cloud is rendered at one
sample per pixel with the network (MRPNN) and with plain delta tracking at
1, 2, 4, ... samples per pixel, all against a 1024-spp delta-tracking
reference of the same scene.  The network's single sample must be worth
several path-tracing samples: the smallest delta-tracking sample count that
reaches its RMSE is the gain.

Measured: delta tracking needs 32-64 spp to match one network sample (RMSE
0.049 with the shipped weights, 0.061 after the short fine-tune; 1-spp delta
tracking 0.295).  Bounds: the fine-tuned network beats 16-spp and the shipped
one 32-spp delta tracking.  A short fine-tune on 4096 noisy 16-spp targets
moves the weights towards that noise, so it may cost image quality against the
shipped weights; it must stay within 1.5x of them."""
import sys

import numpy as np
import pytest

import internship_amd as ia
from internship_amd import Network, Trainer, VolumeRender, camera_basis, datagen, generate_cloud

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def test_gpu_trained_network_gains_over_path_tracing():
    _say("case: fine-tune on 2 clouds x 2048 points (16-spp targets), 24 Adam steps of 1024")
    specs = datagen.random_specs(2, seed=300)
    feat, aux, _ = datagen.make_dataset(specs, res=64, n_per_cloud=2048, spp=16, multi=32, seed=3)
    w0 = ia.load_reference_weights()
    tr = Trainer(w0, max_batch=1024)
    try:
        fb, ab = ia.DeviceBuffer(feat.nbytes), ia.DeviceBuffer(aux.nbytes)
        fb.upload(np.ascontiguousarray(feat))
        ab.upload(np.ascontiguousarray(aux))
        n_batches = len(feat) // 1024
        losses = [tr.step(datagen.DeviceView(fb, (i % n_batches) * 1024 * 576 * 4),
                          datagen.DeviceView(ab, (i % n_batches) * 1024 * 16), 1024, 1e-4) for i in range(24)]
        w1 = tr.weights()
    finally:
        tr.close()
    assert np.all(np.isfinite(losses)) and np.all(np.isfinite(w1))
    _say("training loss %.4g -> %.4g" % (losses[0], losses[-1]))

    _say("case: held-out cloud 64^3 seed 350, 96x96, alpha 128, albedo 0.99, 128 bounces")
    vol = VolumeRender(64)
    try:
        vol.set_datas(generate_cloud(64, seed=350))
        vol.update()
        vol.set_scatter_rate((0.99, 0.99, 0.99))
        ori = 2.5 * np.array([0.75, 0.75, 0.75]) / np.linalg.norm([0.75, 0.75, 0.75])
        right, up = camera_basis(ori, 40.0, 1.0)
        light = np.array([0.0, 1.0, 0.3]) / np.linalg.norm([0.0, 1.0, 0.3])
        kw = dict(g=0.857, alpha=128.0, multi_scatter=128)
        ref = vol.render(96, ori, up, right, light, rt=ia.PT, sample_num=1024, seed=1000, **kw)
        rm = {}
        for name, w in (("shipped", w0), ("trained", w1)):
            net = Network(w)
            vol.set_network(net)
            img = vol.render(96, ori, up, right, light, rt=ia.MRPNN, sample_num=1, seed=7, mlp_precision=1, **kw)
            vol.set_network(None)
            net.close()
            rm[name] = _rmse(img, ref)
        dt = {n: _rmse(vol.render(96, ori, up, right, light, rt=ia.PT, sample_num=n, seed=2000, **kw), ref)
              for n in (1, 2, 4, 8, 16, 32, 64)}
    finally:
        vol.close()
    n_eq = min([n for n, r in dt.items() if r <= rm["trained"]] + [128])
    _say("1-spp RMSE: network shipped %.5g, trained %.5g; delta tracking %s; equal RMSE at %d spp"
         % (rm["shipped"], rm["trained"], {n: round(r, 5) for n, r in dt.items()}, n_eq))
    assert rm["trained"] <= dt[16]
    assert rm["shipped"] <= dt[32]
    assert n_eq >= 32
    assert rm["trained"] <= 1.5 * rm["shipped"]
