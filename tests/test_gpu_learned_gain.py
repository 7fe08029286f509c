"""The learned part, end to end on the GPU, at a reduced config-4 recipe, and
what it gains over the method without it.

Data: Perlin-fBm clouds (seeds 100..107, 64^3, thresholds U[0.3,0.5]) x 2 views
(extinction U[64,512], sun on the upper hemisphere, g U[0.7,0.9]), points
sampled as `MeanFreePathSample` does (dataGen.cu:174-216), the renderer's
gather, `GetSample` targets (64 spp, 128 bounces, albedo 0.999); the last
cloud is held out.  Training starts from seeded random weights (the
uniform +-1/sqrt(fan_in) default of the reference trainer's layers) on the
BF16 matrix-core trainer with Adam (lr 1e-3).

Checks:
- at the random initial weights, the BF16 trainer's gradient on a batch of
  the generated samples matches the oracle's BF16-operand gradient
  (relative L2 < 1e-5, the bar of test_gpu_train.py);
- training lowers the held-out loss to below a third of its initial value,
  and its held-out relative L2 is no worse than the shipped weights' on the
  same samples;
- on a held-out scene (cloud seed 350) one network sample per pixel with the
  trained weights is worth more than 16 samples of plain delta tracking
  against a 1024-spp reference, and is no worse than the shipped network."""
import sys

import numpy as np
import pytest

import internship_amd as ia
import oracle
from internship_amd import Network, Trainer, VolumeRender, camera_basis, datagen, generate_cloud, learn

pytestmark = pytest.mark.gpu

BOUNCES = 128


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


@pytest.fixture(scope="module")
def data():
    _say("case: reduced config-4 data: 8 clouds 64^3 x 2 views x 4096 points, 64-spp targets, %d bounces" % BOUNCES)
    views = datagen.config4_views(8, 100, 2, seed=4)
    ds, st = datagen.make_config4(views, res=64, samples_per_view=4096, spp=64, bounces=BOUNCES, albedo=0.999, seed=5)
    n_train = int(np.argmax(ds.cloud >= 7))
    return ds, n_train


def test_gpu_gradient_at_random_init_matches_oracle(data):
    ds, n_train = data
    _say("case: BF16 trainer gradient at random initial weights, 256 generated samples")
    w0 = learn.init_weights(0, float(ds.aux_host[:n_train, 3].mean()), float(ds.aux_host[0, 2]))
    n = 256
    feat = ds.feat.download(np.float32, (n, 576))
    aux = ds.aux_host[:n]
    tr = Trainer(w0, max_batch=n, precision=1)
    try:
        g, loss = tr.gradient(ds.feat, ds.aux, n)
    finally:
        tr.close()
    og, oloss, _ = oracle.mrpnn_grad(w0, feat, aux, bf16=True)
    rel = np.linalg.norm(g - og) / np.linalg.norm(og)
    _say("gradient rel L2 %.3g, loss %.6g / %.6g" % (rel, loss, oloss))
    assert np.linalg.norm(og) > 0 and rel < 1e-5
    assert abs(loss - oloss) <= 1e-5 * abs(oloss)


def test_gpu_training_from_random_init_gains(data):
    ds, n_train = data
    w0 = learn.init_weights(0, float(ds.aux_host[:n_train, 3].mean()), float(ds.aux_host[0, 2]))
    B = 4096
    tr = Trainer(w0, max_batch=B, precision=1)
    loss0, rel0, _ = learn.evaluate(tr, ds, n_train, ds.n, B)
    tr.close()
    sh = Trainer(ia.load_reference_weights(), max_batch=B, precision=1)
    loss_s, rel_s, _ = learn.evaluate(sh, ds, n_train, ds.n, B)
    sh.close()
    _say("case: 30 epochs of Adam (lr 1e-3, batch %d) from random weights" % B)
    w1, hist, _ = learn.train(ds, n_train, w0, epochs=30, batch=B, lr=1e-3, seed=0, precision=1,
                              eval_range=(n_train, ds.n), eval_every=10)
    loss1, rel1 = hist[-1][2], hist[-1][3]
    _say("held out: loss %.4g -> %.4g (shipped weights %.4g), rel L2 %.4g -> %.4g (shipped %.4g)"
         % (loss0, loss1, loss_s, rel0, rel1, rel_s))
    assert np.all(np.isfinite(w1))
    assert loss1 < loss0 / 3
    assert rel1 <= rel_s

    _say("case: held-out cloud 64^3 seed 350, 96x96, alpha 128, albedo 0.999, g 0.857, %d bounces" % BOUNCES)
    vol = VolumeRender(64)
    try:
        vol.set_datas(generate_cloud(64, seed=350))
        vol.update()
        vol.set_scatter_rate((0.999, 0.999, 0.999))
        ori = 2.5 * np.array([0.75, 0.75, 0.75]) / np.linalg.norm([0.75, 0.75, 0.75])
        right, up = camera_basis(ori, 40.0, 1.0)
        light = np.array([0.0, 1.0, 0.3]) / np.linalg.norm([0.0, 1.0, 0.3])
        kw = dict(g=0.857, alpha=128.0, multi_scatter=BOUNCES)
        ref = vol.render(96, ori, up, right, light, rt=ia.PT, sample_num=1024, seed=1000, **kw)
        rm = {}
        for name, w in (("shipped", ia.load_reference_weights()), ("trained", w1)):
            net = Network(w)
            vol.set_network(net)
            img = vol.render(96, ori, up, right, light, rt=ia.MRPNN, sample_num=1, seed=7, mlp_precision=1, **kw)
            vol.set_network(None)
            net.close()
            rm[name] = _rmse(img, ref)
        dt = {n: _rmse(vol.render(96, ori, up, right, light, rt=ia.PT, sample_num=n, seed=2000, **kw), ref)
              for n in (1, 4, 16, 64)}
    finally:
        vol.close()
    _say("1-spp RMSE: network shipped %.5g, trained %.5g; delta tracking %s"
         % (rm["shipped"], rm["trained"], {n: round(r, 5) for n, r in dt.items()}))
    assert rm["trained"] <= dt[16]
    assert rm["trained"] <= rm["shipped"]
