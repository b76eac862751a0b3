"""CPU pins of the learned-confidence oracle (oracle/learned.py): finite
differences for the float64 gradient, closed forms for the loss and the
features, the BF16 rounding rule, and the BF16-emulated gradient against the
exact one."""
import numpy as np
import pytest

from oracle import learned as L


def _data(n=96, seed=0):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 32), np.float32)
    X[:, :25] = rng.uniform(-0.3, 0.3, (n, 25))
    X[:, 25] = rng.uniform(0.1, 1.0, n)
    X[:, 26] = rng.uniform(0, 1, n)
    r = rng.normal(0, 0.004, n).astype(np.float32)
    return X, r


def test_bf16_rounding_ties_to_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)       # BF16 spacing at 1
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -1 - 2 ** -8], np.float32)
    out = L.bf16(vals)
    assert out[0] == one
    assert out[1] == one                      # tie -> even (1.0)
    assert out[2] == one + 2 * ulp            # tie -> even (1 + 2^-6)
    assert out[3] == one + ulp                # above the tie
    assert out[4] == -one
    bits = (L.bf16(np.array([3.14159], np.float32)).view(np.uint32) >> 16).astype(np.uint16)
    assert np.array_equal(L.bf16_from_bits(bits), L.bf16(np.array([3.14159], np.float32)))


def test_loss_derivative_closed_form():
    o = np.linspace(-6, 6, 41)
    r = np.full_like(o, 0.004)
    eps = 1e-6
    num = (L.loss_terms(o + eps, r)[0] - L.loss_terms(o - eps, r)[0]) / (2 * eps)
    assert np.allclose(L.loss_terms(o, r)[1], num, rtol=1e-6, atol=1e-9)


def test_loss_minimiser_is_inverse_variance():
    # constant output: mean dl/do = 0 at c = s^2 / E[r^2]
    rng = np.random.default_rng(3)
    r = rng.normal(0, 0.006, 200000)
    c_star = L.RES_SCALE ** 2 / np.mean(r ** 2)
    o_star = np.log(c_star / (1 - c_star))
    assert abs(np.mean(L.loss_terms(np.full_like(r, o_star), r)[1])) < 1e-12
    # and the loss there is below that of nearby outputs
    l0 = np.mean(L.loss_terms(np.full_like(r, o_star), r)[0])
    for d in (-0.3, 0.3):
        assert np.mean(L.loss_terms(np.full_like(r, o_star + d), r)[0]) > l0


def test_grads64_match_finite_differences():
    X, r = _data()
    p = L.init_params(1).astype(np.float64)
    p[L.OFF["b4"]] = -1.0
    g = L.grads64(p, X, r)
    rng = np.random.default_rng(5)
    idx = []
    for k, sh in L.SHAPES.items():
        n = int(np.prod(sh))
        idx += list(L.OFF[k] + rng.choice(n, size=min(n, 8), replace=False))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (L.loss64(pp, X, r) - L.loss64(pm, X, r)) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-9, (i, num, g[i])


def test_bf16_grads_track_exact_grads():
    X, r = _data(2048, seed=7)
    Xb = L.bf16(X)
    p = L.init_params(2)
    g64 = L.grads64(p.astype(np.float64), Xb, r)
    gb, loss = L.grads_bf16(p, Xb, r)
    assert np.isfinite(loss)
    for k, sh in L.SHAPES.items():
        a = g64[L.OFF[k]:L.OFF[k] + int(np.prod(sh))]
        b = gb[L.OFF[k]:L.OFF[k] + int(np.prod(sh))]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-12)
        assert rel < 0.03, (k, rel)
    # a dropped ReLU derivative would not pass
    F = L.forward_bf16(p, Xb)
    assert (F["z2"] <= 0).mean() > 0.2


def test_features_fronto_parallel_plane():
    H, W = 24, 32
    z = 1.5
    K = (525.0, 525.0, 15.5, 11.5)
    depth = np.full((H, W), int(round(z * 5000)), np.uint16)
    F, valid = L.features(depth, K, 1 / 5000.0, 4.0)
    assert valid.all()
    inner = F[2:-2, 2:-2]
    assert np.all(inner[:, :, :25] == 0.0)
    assert np.allclose(F[:, :, 25], np.float32(depth[0, 0]) * np.float32(1 / 5000.0) * 0.25)
    # the normal is the optical axis: cos = z / |ray| = 1 / sqrt(u^2 + v^2 + 1)
    ys, xs = np.mgrid[1:H - 1, 1:W - 1]
    u = (xs - K[2]) / K[0]
    v = (ys - K[3]) / K[1]
    assert np.allclose(F[1:-1, 1:-1, 26], 1 / np.sqrt(u * u + v * v + 1), atol=2e-6)
    # border pixels miss a neighbour: cos 0; patch entries outside the image: -1
    assert np.all(F[0, :, 26] == 0.0)
    assert np.all(F[0, 5, 0:5] == -1.0)


def test_features_tilted_plane_incidence():
    # plane through (0, 0, 2) with normal n: depth along each ray from the plane equation
    H, W = 40, 48
    K = (300.0, 300.0, 23.5, 19.5)
    n = np.array([0.3, -0.2, 1.0])
    n /= np.linalg.norm(n)
    ys, xs = np.mgrid[0:H, 0:W]
    rays = np.stack([(xs - K[2]) / K[0], (ys - K[3]) / K[1], np.ones_like(xs, float)], -1)
    z = (n @ np.array([0, 0, 2.0])) / (rays @ n)
    depth = np.round(z * 5000).astype(np.uint16)
    F, valid = L.features(depth, K, 1 / 5000.0, 4.0)
    cos_true = np.abs(rays @ n) / np.linalg.norm(rays, axis=-1)
    # u16 depth quantisation (0.2 mm) perturbs the finite-difference normal
    assert np.abs(F[1:-1, 1:-1, 26] - cos_true[1:-1, 1:-1]).max() < 0.05


def test_features_depth_edge_and_missing():
    H, W = 16, 16
    depth = np.full((H, W), 10000, np.uint16)   # 2 m
    depth[:, 8:] = 15000                        # 3 m behind
    depth[4, 4] = 0                             # missing
    F, valid = L.features(depth, (500.0, 500.0, 7.5, 7.5), 1 / 5000.0, 4.0)
    assert not valid[4, 4] and np.all(F[4, 4] == 0)
    # pixel (6, 7): right neighbours at dx=1,2 are the background: 3/2 - 1 = 0.5
    f = F[6, 7, :25].reshape(5, 5)
    assert np.allclose(f[:, 3:], 0.5) and np.allclose(f[:, :3], 0.0)
    # pixel (6, 8): left neighbours are the foreground: 2/3 - 1
    f = F[6, 8, :25].reshape(5, 5)
    assert np.allclose(f[:, :2], np.float32(2 / 3) - 1, atol=1e-7)


def _oracle_frames(n=3):
    from fastfusion_amd import synth
    sc = synth.Scene.room_scene(7, (4.0, 4.0, 3.0), 3, 3, 0.1, 0.4)
    K = (262.5, 262.5, 159.5, 119.5)
    R, t = sc.walk(n, 8, 0.5, 3.0, 0.02, 2.0)
    return K, R, t, [sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, seed=40 + i) for i in range(n)]


def _state(o):
    n = o.counts()["n_leaves_used"]
    return o.tree().copy(), o.bricks(np.arange(n))


def test_oracle_confidence_hook():
    """The confidence multiplies the weight increment only: c = 1 is the plain
    update bit for bit; c = 1/2 on the first frame halves every weight exactly
    and leaves distances and colors unchanged (power-of-two scaling is exact in
    round-toward-zero products and in RCPPS).  Applying c to the distance, to
    the accumulated weight or to the colour increment fails one of these."""
    from oracle.pyoracle import Oracle
    K, R, t, fr = _oracle_frames()
    plain, ones, half = (Oracle(0.01, 0.03, True, 1 << 16, 1 << 20) for _ in range(3))
    for i, (d, bgr) in enumerate(fr):
        plain.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0)
        ones.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0, conf=np.ones(d.shape, np.float32))
        if i == 0:
            half.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0, conf=np.full(d.shape, 0.5, np.float32))
            (ta, (da, wa, ca)), (th, (dh, wh, ch)) = _state(plain), _state(half)
            assert np.array_equal(ta, th)
            assert np.array_equal(da.view(np.uint32), dh.view(np.uint32))
            assert np.array_equal(ca, ch)
            assert np.array_equal((wa * np.float32(0.5)).view(np.uint32), wh.view(np.uint32))
            assert (wa > 0).sum() > 10000
    (ta, (da, wa, ca)), (tb, (db, wb, cb)) = _state(plain), _state(ones)
    assert np.array_equal(ta, tb)
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
    assert np.array_equal(ca, cb)
