"""CPU oracle of the learned per-pixel confidence -- test infrastructure only.

Used by tests/ (and nowhere in the product path) to check the HIP kernels of
fastfusion_amd/csrc/ff_learned.hip.  The reference has no learned component:
its weight is the fixed linear fall-off of getDistanceWeight_AoS_variables_h
(inline_functions.hpp:83-124); the confidence c(pixel) multiplies it.

Three restatements:
  * features(): the 27 per-pixel inputs, float32 in the kernels' order of
    operations (5x5 patch z_k * (1 / z_c) - 1 clamped to [-1, 1], z_c / 4,
    incidence cosine from central differences with u = (x - cx) * (1 / fx));
  * forward_bf16() / grads_bf16(): the network and the gradient of the mean
    loss with the kernels' BF16 rounding points (weights, activations and
    back-propagated deltas rounded to BF16, sums in float32);
  * forward64() / grads64(): the same mathematics in float64 without
    rounding -- pinned by finite differences in tests/test_cpu_learned.py.
Loss per sample, residual r, s = RES_SCALE:  l = c r^2 / (2 s^2) - log(c) / 2.
"""
import numpy as np

NPARAMS = 10177
RES_SCALE = 0.0025
GATE = 0.002  # inference: confidences below it are 0
NIN, NH = 27, 64
OFF = {"W1": 0, "b1": 64 * 27, "W2": 64 * 27 + 64, "b2": 64 * 27 + 64 + 4096,
       "W3": 64 * 27 + 128 + 4096, "b3": 64 * 27 + 128 + 8192, "w4": 64 * 27 + 192 + 8192,
       "b4": 64 * 27 + 256 + 8192}
SHAPES = {"W1": (64, 27), "b1": (64,), "W2": (64, 64), "b2": (64,), "W3": (64, 64), "b3": (64,),
          "w4": (64,), "b4": (1,)}


def unpack(p):
    p = np.asarray(p)
    return {k: p[OFF[k]:OFF[k] + int(np.prod(SHAPES[k]))].reshape(SHAPES[k]) for k in OFF}


def pack(d, dtype=np.float32):
    out = np.zeros(NPARAMS, dtype)
    for k in OFF:
        out[OFF[k]:OFF[k] + int(np.prod(SHAPES[k]))] = np.asarray(d[k]).ravel()
    return out


def init_params(seed=0):
    """He-style initialisation; b4 starts the confidence near 0.5."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / 27), (64, 27)), "b1": np.zeros(64),
         "W2": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b2": np.zeros(64),
         "W3": rng.normal(0, np.sqrt(2 / 64), (64, 64)), "b3": np.zeros(64),
         "w4": rng.normal(0, np.sqrt(1 / 64), 64), "b4": np.zeros(1)}
    return pack(d)


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_from_bits(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def _f(x):
    return np.float32(x)


def features(depth, K, scaling, maxcam):
    """(H, W, 32) float32 features and the (H, W) valid mask."""
    depth = np.asarray(depth, np.uint16)
    H, W = depth.shape
    fx, fy, cx, cy = (np.float32(k) for k in K)
    ifx, ify = np.float32(1) / fx, np.float32(1) / fy
    z = depth.astype(np.float32) * np.float32(scaling)
    zp = np.zeros((H + 4, W + 4), np.float32)
    zp[2:H + 2, 2:W + 2] = z
    valid = (z > 0) & (z < np.float32(maxcam))
    zc = np.where(valid, z, np.float32(1.0))
    rc = np.float32(1.0) / zc
    F = np.zeros((H, W, 32), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(25):
            dy, dx = k // 5 - 2, k % 5 - 2
            zk = zp[2 + dy:2 + dy + H, 2 + dx:2 + dx + W]
            F[:, :, k] = np.clip(zk * rc - np.float32(1.0), np.float32(-1.0), np.float32(1.0))
        F[:, :, 25] = z * np.float32(0.25)
        xs = np.arange(W, dtype=np.float32)
        ys = np.arange(H, dtype=np.float32)
        uc = ((xs - cx) * ifx)[None, :]
        ul = ((xs - np.float32(1) - cx) * ifx)[None, :]
        ur = ((xs + np.float32(1) - cx) * ifx)[None, :]
        vc = ((ys - cy) * ify)[:, None]
        vu = ((ys - np.float32(1) - cy) * ify)[:, None]
        vd = ((ys + np.float32(1) - cy) * ify)[:, None]
        zl = zp[2:H + 2, 1:W + 1]
        zr = zp[2:H + 2, 3:W + 3]
        zu = zp[1:H + 1, 2:W + 2]
        zd = zp[3:H + 3, 2:W + 2]
        tx0 = ur * zr - ul * zl
        tx1 = vc * zr - vc * zl
        tx2 = zr - zl
        ty0 = uc * zd - uc * zu
        ty1 = vd * zd - vu * zu
        ty2 = zd - zu
        n0 = tx1 * ty2 - tx2 * ty1
        n1 = tx2 * ty0 - tx0 * ty2
        n2 = tx0 * ty1 - tx1 * ty0
        r0 = uc * z
        r1 = vc * z
        r2 = z
        nn = (n0 * n0 + n1 * n1) + n2 * n2
        rr = (r0 * r0 + r1 * r1) + r2 * r2
        dot = (n0 * r0 + n1 * r1) + n2 * r2
        cosv = np.abs(dot) / np.sqrt(nn * rr)
        ok = (zl > 0) & (zr > 0) & (zu > 0) & (zd > 0) & (nn > 0) & (rr > 0)
        F[:, :, 26] = np.where(ok, cosv, np.float32(0.0))
    F[~valid] = 0.0
    return F, valid


def forward_bf16(p, X):
    """X: (n, 32) float32 features already rounded to BF16.  Returns a dict of
    the pre-activations z*, the BF16 activations h*, the output o and c."""
    d = unpack(np.asarray(p, np.float32))
    W1, W2, W3 = bf16(d["W1"]), bf16(d["W2"]), bf16(d["W3"])
    w4 = bf16(d["w4"])
    X = np.asarray(X, np.float32)[:, :NIN]
    z1 = (X @ W1.T).astype(np.float32) + d["b1"].astype(np.float32)
    h1 = bf16(np.maximum(z1, 0))
    z2 = (h1 @ W2.T).astype(np.float32) + d["b2"].astype(np.float32)
    h2 = bf16(np.maximum(z2, 0))
    z3 = (h2 @ W3.T).astype(np.float32) + d["b3"].astype(np.float32)
    h3 = bf16(np.maximum(z3, 0))
    o = (h3 @ w4).astype(np.float32) + np.float32(d["b4"][0])
    c = (1.0 / (1.0 + np.exp(-o.astype(np.float64)))).astype(np.float32)
    return {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "h3": h3, "o": o, "c": c,
            "W1": W1, "W2": W2, "W3": W3, "w4": w4}


def loss_terms(o, r, s=RES_SCALE):
    o = np.asarray(o, np.float64)
    c = 1.0 / (1.0 + np.exp(-o))
    logc = -np.logaddexp(0.0, -o)
    r2s = 0.5 * np.asarray(r, np.float64) ** 2 / (s * s)
    return c * r2s - 0.5 * logc, (1.0 - c) * (c * r2s - 0.5)


def grads_bf16(p, X, r):
    """Gradient of the mean loss, with the kernels' BF16 rounding points."""
    F = forward_bf16(p, X)
    n = len(r)
    X = np.asarray(X, np.float32)[:, :NIN]
    r = np.asarray(r, np.float32)
    r2s = (np.float32(0.5) * r * r * np.float32(1.0 / (RES_SCALE * RES_SCALE))).astype(np.float32)
    c = F["c"]
    go = ((np.float32(1) - c) * (c * r2s - np.float32(0.5)) * np.float32(1.0 / n)).astype(np.float32)
    loss = float(np.mean(loss_terms(F["o"], r)[0]))
    d3 = bf16(np.where(F["z3"] > 0, go[:, None] * F["w4"][None, :], 0).astype(np.float32))
    d2 = bf16(np.where(F["z2"] > 0, d3 @ F["W3"], 0).astype(np.float32))
    d1 = bf16(np.where(F["z1"] > 0, d2 @ F["W2"], 0).astype(np.float32))
    g = {"W3": d3.T @ F["h2"], "W2": d2.T @ F["h1"], "W1": d1.T @ X,
         "b3": d3.sum(0), "b2": d2.sum(0), "b1": d1.sum(0),
         "w4": F["h3"].T @ go, "b4": np.array([go.sum()], np.float32)}
    return pack(g), loss


def forward64(p, X):
    d = unpack(np.asarray(p, np.float64))
    X = np.asarray(X, np.float64)[:, :NIN]
    z1 = X @ d["W1"].T + d["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ d["W3"].T + d["b3"]
    h3 = np.maximum(z3, 0)
    o = h3 @ d["w4"] + d["b4"][0]
    return {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "h3": h3, "o": o, "d": d}


def loss64(p, X, r):
    return float(np.mean(loss_terms(forward64(p, X)["o"], r)[0]))


def grads64(p, X, r):
    F = forward64(p, X)
    d = F["d"]
    n = len(r)
    X = np.asarray(X, np.float64)[:, :NIN]
    go = loss_terms(F["o"], r)[1] / n
    d3 = np.where(F["z3"] > 0, go[:, None] * d["w4"][None, :], 0)
    d2 = np.where(F["z2"] > 0, d3 @ d["W3"], 0)
    d1 = np.where(F["z1"] > 0, d2 @ d["W2"], 0)
    g = {"W3": d3.T @ F["h2"], "W2": d2.T @ F["h1"], "W1": d1.T @ X,
         "b3": d3.sum(0), "b2": d2.sum(0), "b1": d1.sum(0),
         "w4": F["h3"].T @ go, "b4": np.array([go.sum()])}
    return pack(g, np.float64)
