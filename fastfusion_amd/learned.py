"""Learned per-pixel confidence: data, training and evaluation drivers.

The network and its training run in the HIP library (ff_learned.hip, exposed
as ``FusionMipMapGPU.setLearnedWeighting`` and ``Trainer``); this module
builds the synthetic training set and drives it.

Training data (one room per seed): the room is ray-cast along a seeded
random walk with heavier sensor noise (sigma x2, 3 % flying pixels at depth
edges, 2 % dropout), all frames are fused with the plain method, and for
every ``collect_every``-th frame the residuals of its pixels against that
converged map become samples.  Evaluation fuses held-out rooms with plain and
with learned weighting and compares the RMS of the analytic scene SDF at the
final mesh's vertices.

    python -m fastfusion_amd.learned train --out fastfusion_amd/weights/confidence_mlp.npy
    python -m fastfusion_amd.learned eval --weights fastfusion_amd/weights/confidence_mlp.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import FusionMipMapGPU, MLP_NPARAMS, Trainer, synth

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "confidence_mlp.npy")

# heavier-noise regime of the learned configuration, 640x480, 5 mm voxels
CFG = dict(scale=0.005, threshold=0.015, maxcam=4.0, w=640, h=480, noise=2.0, flying=0.03, dropout=0.02)
TRAIN_SEEDS = range(1000, 1050)
EVAL_SEEDS = range(2000, 2010)


def load_weights(path=WEIGHTS):
    p = np.load(path, allow_pickle=False).astype(np.float32).ravel()
    if p.size != MLP_NPARAMS:
        raise ValueError(f"{path}: expected {MLP_NPARAMS} parameters, found {p.size}")
    return p


def init_params(seed=0):
    """He-style initialisation; the output bias starts the confidence at 0.5."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 27), 64 * 27), np.zeros(64),
             rng.normal(0, np.sqrt(2 / 64), 4096), np.zeros(64),
             rng.normal(0, np.sqrt(2 / 64), 4096), np.zeros(64),
             rng.normal(0, np.sqrt(1 / 64), 64), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def room(seed):
    sc = synth.Scene.room_scene(seed, (4.0, 4.0, 3.0), 6, 4, 0.1, 0.5)
    return sc


def frames(sc, n, seed, cfg=CFG):
    R, t = sc.walk(n, seed + 7, 0.5, 3.0, 0.02, 2.0)
    for i in range(n):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], synth.K_DEFAULT, cfg["w"], cfg["h"], noise=cfg["noise"],
                           dropout=cfg["dropout"], flying=cfg["flying"], seed=seed * 100003 + i)
        yield i, d, bgr, R[i], t[i]


def fuse(sc, n, seed, params=None, cfg=CFG, max_leaves=1 << 20, keep_frames=None):
    """Fuses n frames of room sc; learned weighting when params is given.
    keep_frames: optional set of frame indices whose (depth, R, t) are returned."""
    fz = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=max_leaves)
    fz.setIncrementalMeshing(False)
    if params is not None:
        fz.setLearnedWeighting(params)
    kept = []
    for i, d, bgr, R, t in frames(sc, n, seed, cfg):
        fz.addMap(d, R, t, synth.K_DEFAULT, bgr, 1 / 5000.0, cfg["maxcam"])
        if keep_frames is not None and i in keep_frames:
            kept.append((d, R, t))
    return fz, kept


def mesh_errors(fz, sc):
    """Analytic scene SDF at the vertices of the map's full mesh."""
    fz.setIncrementalMeshing(True)
    fz.updateMeshes()
    m = fz.getMeshInterleavedMarchingCubes()
    return sc.sdf(m.vertices).astype(np.float64) if len(m.vertices) else np.zeros(0)


def rms_error(fz, sc):
    e = mesh_errors(fz, sc)
    return (float(np.sqrt(np.mean(e * e))) if len(e) else float("nan")), len(e)


def collect(trainer, seeds, n_frames, every, keep, cfg=CFG, log=print):
    for k, s in enumerate(seeds):
        t0 = time.perf_counter()
        sc = room(s)
        want = set(range(every // 2, n_frames, every))
        fz, kept = fuse(sc, n_frames, s, None, cfg, keep_frames=want)
        added = 0
        for d, R, t in kept:
            added += trainer.collect(fz, d, R, t, synth.K_DEFAULT, 1 / 5000.0, cfg["maxcam"], keep)
        fz.close()
        log(f"[learned] room {s}: {added} samples ({trainer.size()} total) in {time.perf_counter() - t0:.1f} s")


def train(args):
    log = lambda m: print(m, flush=True)  # noqa: E731
    tr = Trainer(init_params(args.seed), capacity=args.capacity, seed=args.seed)
    collect(tr, range(args.seed_base, args.seed_base + args.rooms), args.frames, args.collect_every, args.keep,
            log=log)
    t0 = time.perf_counter()
    done = 0
    chunk = 1000
    while done < args.steps:
        n = min(chunk, args.steps - done)
        lr = args.lr if done < int(0.75 * args.steps) else args.lr * 0.2
        loss = tr.train(n, args.batch, lr)
        done += n
        log(f"[learned] step {done}: loss {loss:.5f} ({time.perf_counter() - t0:.1f} s)")
    p = tr.params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, p)
    log(f"[learned] wrote {args.out}")


def evaluate(params, seeds, n_frames, cfg=CFG, log=print):
    rows = []
    for s in seeds:
        sc = room(s)
        out = {"seed": s}
        for name, p in (("plain", None), ("learned", params)):
            fz, _ = fuse(sc, n_frames, s, p, cfg)
            e = mesh_errors(fz, sc)
            fz.close()
            out[name] = float(np.sqrt(np.mean(e * e))) if len(e) else float("nan")
            out[name + "_vertices"] = int(len(e))
            # accuracy of the vertices within 2 cm of a surface, and the share beyond 5 cm
            inl = np.abs(e) < 0.02
            out[name + "_rms_within_2cm"] = float(np.sqrt(np.mean(e[inl] ** 2))) if inl.any() else float("nan")
            out[name + "_frac_beyond_5cm"] = float(np.mean(np.abs(e) > 0.05)) if len(e) else 0.0
        rows.append(out)
        log(f"[learned] room {s}: plain {out['plain'] * 1e3:.3f} mm, learned {out['learned'] * 1e3:.3f} mm")
    plain = float(np.sqrt(np.mean([r["plain"] ** 2 for r in rows])))
    learned = float(np.sqrt(np.mean([r["learned"] ** 2 for r in rows])))
    vp = sum(r["plain_vertices"] for r in rows)
    vl = sum(r["learned_vertices"] for r in rows)
    return {"rooms": rows, "rms_plain_m": plain, "rms_learned_m": learned,
            "reduction": 1.0 - learned / plain, "vertex_ratio": vl / max(vp, 1)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("train")
    a.add_argument("--out", default=WEIGHTS)
    a.add_argument("--rooms", type=int, default=50)
    a.add_argument("--seed-base", type=int, default=TRAIN_SEEDS.start)
    a.add_argument("--frames", type=int, default=200)
    a.add_argument("--collect-every", type=int, default=20)
    a.add_argument("--keep", type=float, default=0.05)
    a.add_argument("--capacity", type=int, default=1 << 23)
    a.add_argument("--steps", type=int, default=20000)
    a.add_argument("--batch", type=int, default=65536)
    a.add_argument("--lr", type=float, default=3e-3)
    a.add_argument("--seed", type=int, default=1)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--rooms", type=int, default=10)
    e.add_argument("--seed-base", type=int, default=EVAL_SEEDS.start)
    e.add_argument("--frames", type=int, default=200)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        train(args)
    else:
        res = evaluate(load_weights(args.weights), range(args.seed_base, args.seed_base + args.rooms), args.frames,
                       log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
