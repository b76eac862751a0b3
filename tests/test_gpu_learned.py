"""GPU checks of the learned confidence against the CPU oracle
(oracle/learned.py): per-pixel inference (features + BF16 MFMA network),
the SDF update with the confidence factor (bit-exact against the C oracle fed
the same confidence image), the training kernels' gradient, and that training
lowers the loss toward the inverse-variance optimum."""
import os

import numpy as np
import pytest

from fastfusion_amd import FusionMipMapGPU, Trainer, learned, synth
from oracle import learned as L
from tests.helpers import compare_state, make_pair

pytestmark = pytest.mark.gpu


def _frame(seed=3, w=320, h=240):
    sc = synth.Scene.room_scene(seed, (4.0, 4.0, 3.0), 6, 4, 0.1, 0.5)
    R, t = sc.walk(4, seed + 1, 0.5, 3.0, 0.02, 2.0)
    K = (262.5, 262.5, 159.5, 119.5)
    fr = [sc.render(R[i].reshape(3, 3), t[i], K, w, h, noise=2.0, dropout=0.02, flying=0.03, seed=50 + i)
          for i in range(len(R))]
    return sc, R, t, K, fr


@pytest.mark.parametrize("frame", [0, 1, 3])
def test_confidence_matches_oracle(frame):
    """Per-pixel confidences of the first frame (run with the frame) and of
    later ones (run beside the frame's allocation, after its images are up)."""
    sc, R, t, K, fr = _frame()
    p = L.init_params(11)
    p[L.OFF["b4"]] = 0.3
    fz = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 16)
    fz.setLearnedWeighting(p)
    for i in range(frame + 1):
        d, bgr = fr[i]
        fz.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0)
        fz.updateMeshes()
    c_gpu = fz.confidence(d.shape[1], d.shape[0])
    F, valid = L.features(d, K, 1 / 5000.0, 4.0)
    c_ora = np.zeros(d.shape, np.float32)
    c_ora[valid] = L.forward_bf16(p, L.bf16(F[valid]))["c"]
    assert np.all(c_gpu[~valid] == 0.0)
    # gate: confidences under it are 0 (pixels within rounding of the gate excluded)
    gated = valid & (c_ora < L.GATE * 0.999)
    assert np.all(c_gpu[gated] == 0.0)
    c_ora[gated] = 0.0
    near = np.abs(c_ora - L.GATE) < L.GATE * 1e-3
    err = np.abs(c_gpu - c_ora)[valid & ~near]
    assert err.max() < 2e-2 and err.mean() < 1e-3, (err.max(), err.mean())
    assert c_ora[valid].std() > 0.01  # the network output actually varies


def test_sdf_with_confidence_bit_exact():
    sc, R, t, K, fr = _frame(5)
    p = L.init_params(12)
    gpu, ora = make_pair(0.01, 0.03, max_leaves=1 << 16)
    gpu.setLearnedWeighting(p)
    for i, (d, bgr) in enumerate(fr):
        nq_g = gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0)
        conf = gpu.confidence(d.shape[1], d.shape[0])
        nq_o = ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0, conf=conf)
        assert nq_g == nq_o
        compare_state(gpu, ora, d.shape[1], d.shape[0], tag=f"frame {i}")


def _dataset(n, seed=0):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 32), np.float32)
    X[:, :25] = rng.uniform(-0.2, 0.2, (n, 25))
    X[:, 25] = rng.uniform(0.1, 1.0, n)
    X[:, 26] = rng.uniform(0, 1, n)
    Xb = L.bf16(X)
    bits = (Xb.view(np.uint32) >> 16).astype(np.uint16)
    # noise grows with depth feature: sigma = 2 mm .. 20 mm
    sigma = 0.002 + 0.02 * X[:, 25] ** 2
    r = (rng.normal(0, 1, n) * sigma).astype(np.float32)
    return Xb, bits, r, sigma


def test_training_gradient_matches_oracle():
    Xb, bits, r, _ = _dataset(8192, 1)
    p = L.init_params(13)
    p[L.OFF["b4"]] = -0.5
    tr = Trainer(p, capacity=8192, seed=3)
    tr.set_samples(bits, r)
    idx = np.random.default_rng(2).choice(8192, 4096, replace=False).astype(np.uint32)
    g_gpu, loss_gpu = tr.grad(idx)
    g_ora, loss_ora = L.grads_bf16(p, Xb[idx], r[idx])
    assert abs(loss_gpu - loss_ora) <= 1e-3 * abs(loss_ora) + 1e-6, (loss_gpu, loss_ora)
    for k, sh in L.SHAPES.items():
        a = g_ora[L.OFF[k]:L.OFF[k] + int(np.prod(sh))]
        b = g_gpu[L.OFF[k]:L.OFF[k] + int(np.prod(sh))]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-12)
        assert rel < 1e-2, (k, rel)
    tr.close()


def test_training_reaches_inverse_variance():
    Xb, bits, r, sigma = _dataset(1 << 17, 4)
    tr = Trainer(L.init_params(14), capacity=1 << 17, seed=5)
    tr.set_samples(bits, r)
    l0 = tr.grad(np.arange(8192, dtype=np.uint32))[1]
    loss = tr.train(600, 16384, 2e-3)
    assert loss < l0 - 0.1, (l0, loss)
    # learned c against s^2 / sigma^2 (capped at 1): correlated in log space
    F = L.forward_bf16(tr.params(), Xb[:8192])
    c_star = np.minimum(L.RES_SCALE ** 2 / sigma[:8192] ** 2, 1.0)
    corr = np.corrcoef(np.log(F["c"]), np.log(c_star))[0, 1]
    assert corr > 0.9, corr
    tr.close()


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS), reason="no trained weights committed")
def test_trained_weights_lower_surface_error():
    p = learned.load_weights()
    res = learned.evaluate(p, [2000, 2001], 80)
    assert res["rms_learned_m"] <= 0.85 * res["rms_plain_m"], res
