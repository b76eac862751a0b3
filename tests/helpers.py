"""Shared helpers of the parity tests: run the HIP library and the oracle on
the same synthetic frames and compare their state bit for bit."""
import numpy as np

from fastfusion_amd import FusionMipMapGPU, synth
from oracle.pyoracle import Oracle

K = synth.K_DEFAULT


def make_pair(scale, threshold, color=True, max_leaves=1 << 17, max_branches=1 << 22):
    gpu = FusionMipMapGPU(0, 0, 0, scale, threshold, 0, color, max_leaves=max_leaves, max_branches=max_branches)
    ora = Oracle(scale, threshold, color, max_leaves=max_leaves, max_branches=max_branches)
    return gpu, ora


def compare_state(gpu, ora, w, h, check_bricks=True, tag=""):
    cg, co = gpu.counts(), ora.counts()
    for k in ("n_leaves_used", "n_branches_used", "n_queued", "side", "grows"):
        assert cg[k] == co[k], f"{tag} {k}: gpu {cg[k]} oracle {co[k]}"
    assert np.array_equal(cg["offset"], co["offset"]), f"{tag} offset {cg['offset']} vs {co['offset']}"
    pg, po = gpu.processed(w, h), ora.processed(w, h)
    assert np.array_equal(pg, po), f"{tag} de-dup differs at {np.argwhere(pg != po)[:5]}"
    tg, to = gpu.tree(), ora.tree()
    assert np.array_equal(tg, to), f"{tag} tree differs at {np.flatnonzero(tg != to)[:8]}"
    pos_g, sc_g, par_g = gpu.leaves()
    pos_o, sc_o, par_o = ora.leaves()
    assert np.array_equal(pos_g, pos_o), f"{tag} leaf positions differ"
    assert np.array_equal(sc_g, sc_o), f"{tag} leaf scales differ"
    assert np.array_equal(par_g, par_o), f"{tag} leaf parents differ at {np.flatnonzero(par_g != par_o)[:8]}"
    qg, qo = gpu.queue(), np.sort(ora.queue())
    assert np.array_equal(qg, qo), f"{tag} queued sets differ"
    if check_bricks and len(qg):
        dg, wg, cgc = gpu.bricks(qg)
        do, wo, coc = ora.bricks(qg)
        bad_w = np.flatnonzero(wg.view(np.uint32).ravel() != wo.view(np.uint32).ravel())
        assert bad_w.size == 0, f"{tag} weights differ in {bad_w.size} voxels, first {bad_w[:4]}"
        bad_d = np.flatnonzero(dg.view(np.uint32).ravel() != do.view(np.uint32).ravel())
        assert bad_d.size == 0, (f"{tag} distances differ in {bad_d.size} voxels: "
                                 f"{dg.ravel()[bad_d[:4]]} vs {do.ravel()[bad_d[:4]]}")
        if gpu.color:
            assert np.array_equal(cgc, coc), f"{tag} colors differ in {(cgc != coc).sum()} values"
    return cg


def run_pair(sc, R, t, frames, cfg, gpu, ora, seed0=1000, check_every=1, check_bricks=True, learned=False,
             on_frame=None):
    """Fuses the given frames into both maps (the camera is cfg["K"], else
    this module's K) and compares their state every check_every frames and
    after the last.  learned: the GPU's per-pixel confidence of each frame is
    handed to the oracle.  Returns the queued count of every frame."""
    w, h = cfg["w"], cfg["h"]
    Kc = cfg.get("K", K)
    queued = []
    for k, i in enumerate(frames):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], Kc, w, h, noise=cfg.get("noise", 1.0),
                           dropout=cfg.get("dropout", 0.005), seed=seed0 + i)
        ng = gpu.addMap(d, R[i], t[i], Kc, bgr, 1 / 5000.0, cfg["maxcam"])
        conf = gpu.confidence(w, h) if learned else None
        no = ora.add_map(d, R[i], t[i], Kc, bgr, 1 / 5000.0, cfg["maxcam"], conf=conf)
        assert ng == no, f"frame {i}: gpu queued {ng}, oracle {no}"
        queued.append(ng)
        if (k % check_every) == 0 or k == len(frames) - 1:
            compare_state(gpu, ora, w, h, check_bricks, tag=f"frame {i}")
        if on_frame is not None:
            on_frame(k, i)
    return queued
