#!/usr/bin/env python3
"""Fused + meshed frames/s of the MI355X fusion library on one GPU.

Workload: the 640x480, 5 mm plain-method configuration -- a seeded analytic
4x4x3 m room (walls, 6 axis boxes, 4 spheres), a 1000-frame random walk
0.5-3 m from the surfaces in 2 cm / 2 degree steps, Kinect-style u16 depth
with sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 noise and 0.5 % dropout, BGR8
color, 5 mm voxels, 15 mm truncation, 4 m max depth.  A step is one frame:
addMap (allocation + SDF update) followed by updateMeshes.  All frames are
rendered on the host beforehand and uploaded to HBM before the timed region.

Prints one JSON line.  ``--gpus`` above 1 is not supported (one-GPU build).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fused+meshed frames/sec and HBM GB/s at 640x480, 5mm; RMS surface err vs GT"
HBM_PEAK_GBS = 8000.0


def render_frames(sc, R, t, n, cfg):
    from fastfusion_amd import synth
    w, h = cfg["w"], cfg["h"]
    depth = np.zeros((n, h, w), np.uint16)
    bgr = np.zeros((n, h, w, 3), np.uint8)
    for i in range(n):
        depth[i], bgr[i] = sc.render(R[i].reshape(3, 3), t[i], synth.K_DEFAULT, w, h, noise=cfg["noise"],
                                     dropout=cfg["dropout"], seed=1000 + i)
    return depth, bgr


def cpu_baseline(frame, n_max, R, t, K, cfg, budget_s, what):
    """The oracle (a scalar port of the reference path, 1 core) on the first
    frames of the same workload, bounded by budget_s seconds of oracle time
    (frame(i) gives the host images; producing them is not timed)."""
    from oracle.pyoracle import Oracle
    o = Oracle(cfg["scale"], cfg["threshold"], True, max_leaves=1 << 20, max_branches=1 << 24)
    n = 0
    dt = 0.0
    while n < n_max and dt < budget_s:
        d, bgr = frame(n)
        t0 = time.perf_counter()
        o.add_map(d, R[n], t[n], K, bgr, 1 / 5000.0, cfg["maxcam"])
        o.update_meshes()
        dt += time.perf_counter() - t0
        n += 1
    return {"value": n / dt, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"first {n} frames of the same {what} sequence, oracle addMap + updateMeshes (plain "
                      f"weighting) on 1 host core, {dt:.1f} s"}


LOOP_METRIC = "de-integrated frames/sec (loop closure) at 640x480, 5mm"


def loop_pose(R, t, k):
    """The corrected pose of frame k after a loop closure: a rotation of
    0.1 mrad per frame about z and a drift of 0.2 mm per frame."""
    a = 1e-4 * k
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    return Rz @ np.asarray(R, np.float64).reshape(3, 3), np.asarray(t, np.float64) + 2e-4 * k * np.ones(3)


def loop_cpu_baseline(depth, bgr, R, t, cfg, budget_s):
    """The oracle's de-integration (scalar port, 1 core): the first frames of
    the same walk fused into an oracle map (untimed), then subtracted at their
    corrected poses until budget_s seconds are used."""
    from fastfusion_amd import loopclosure, synth
    from oracle.pyoracle import LoopOracle, Oracle
    o = Oracle(cfg["scale"], cfg["threshold"], True, max_leaves=1 << 20, max_branches=1 << 24)
    queues = []
    t0 = time.perf_counter()
    while len(queues) < len(depth) and time.perf_counter() - t0 < budget_s / 3:
        i = len(queues)
        o.add_map(depth[i], R[i], t[i], synth.K_DEFAULT, bgr[i], 1 / 5000.0, cfg["maxcam"])
        queues.append(np.sort(o.queue()))
    lo = LoopOracle(o, o.counts()["n_leaves_used"])
    n = bricks = 0
    t0 = time.perf_counter()
    while n < len(queues) and time.perf_counter() - t0 < budget_s:
        R2, t2 = loop_pose(R[n], t[n], n)
        lo.delete_frame(loopclosure.depth_metres(depth[n]), loopclosure.planes(bgr[n]), R2, t2, synth.K_DEFAULT,
                        queues[n])
        bricks += len(queues[n])
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "frames/s", "bricks_per_s": bricks / dt, "cores": 1, "kind": "port",
            "sample": f"first {n} frames of the same 640x480 5 mm walk ({bricks} queued bricks) subtracted by "
                      f"the oracle's updateWrapper_deleteloop on 1 host core, {dt:.1f} s"}


def run_loop(args):
    """Loop-closure de-integration: the walk's frames are fused (untimed) with
    their brick queues recorded, then subtracted from a partner map at
    corrected poses, in the reference's pose order, with the depth (metres),
    color planes and queues resident in HBM."""
    from fastfusion_amd import FusionMipMapGPU, hiprt, loopclosure, synth
    n_frames = args.warmup + args.steps
    sc, R, t, cfg = synth.config_c2(n_frames=max(n_frames, 1))
    depth, bgr = render_frames(sc, R, t, n_frames, cfg)
    w, h = cfg["w"], cfg["h"]
    K = synth.K_DEFAULT
    fz = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 21,
                         max_branches=1 << 25)
    d_depth = hiprt.DeviceBuffer.from_array(depth)
    d_bgr = hiprt.DeviceBuffer.from_array(bgr)
    fb, cb = w * h * 2, w * h * 3
    queues = []
    for i in range(n_frames):
        fz.addMapDevice(d_depth.addr(i * fb), d_bgr.addr(i * cb), w, h, R[i], t[i], K, 1 / 5000.0, cfg["maxcam"])
        queues.append(fz.queue())
        if i % 200 == 199:
            print(f"[bench] fused {i + 1} frames", file=sys.stderr, flush=True)
    del d_depth, d_bgr
    counts = fz.counts()
    lp = loopclosure.LoopClosure(fz, counts["n_leaves_used"])
    dm = loopclosure.depth_metres(depth)
    pl = np.ascontiguousarray(np.stack([bgr[..., 2], bgr[..., 1], bgr[..., 0]], 1))  # (n, 3, h, w)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in queues])])
    d_dm = hiprt.DeviceBuffer.from_array(dm)
    d_pl = hiprt.DeviceBuffer.from_array(pl)
    d_q = hiprt.DeviceBuffer.from_array(np.concatenate(queues).astype(np.uint32))
    moved = [loop_pose(R[i], t[i], i) for i in range(n_frames)]
    order = loopclosure.pose_order([m[1] for m in moved], t[:n_frames])
    npix = w * h

    def frame(p):
        R2, t2 = moved[p]
        return (d_dm.addr(p * npix * 4), d_pl.addr(p * 3 * npix), d_pl.addr((p * 3 + 1) * npix),
                d_pl.addr((p * 3 + 2) * npix), w, h, R2, t2, K, d_q.addr(int(qoff[p]) * 4),
                int(qoff[p + 1] - qoff[p]))

    # deleteOldLoopPoses visits at most MAX_NUMBER_POSES_PER_LOOP poses per
    # loop closure: one submission per such group of frames, prepared ahead
    G = loopclosure.MAX_NUMBER_POSES_PER_LOOP

    def groups(poses):
        return [lp.prepare_frames([frame(p) for p in poses[g:g + G]]) for g in range(0, len(poses), G)]

    def submit(prepared):
        done = 0
        for sub in prepared:
            ok, applied = lp.delete_frames(sub)
            assert ok
            done += int(applied.sum())
        return done

    warm, timed = groups(order[:args.warmup]), groups(order[args.warmup:n_frames])
    submit(warm)
    hiprt.synchronize()
    lp.set_timing(True)
    t0 = time.perf_counter()
    bricks = submit(timed)
    hiprt.synchronize()
    dt = time.perf_counter() - t0
    ms, calls, by, units = lp.timing()
    k_ms = ms / max(calls, 1)
    k_bytes = by / max(calls, 1)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        cpu = loop_cpu_baseline(depth, bgr, R, t, cfg, args.cpu_seconds)
    line = {
        "metric": LOOP_METRIC,
        "value": round(args.steps / dt, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded analytic room ray-cast on the host; float depth (metres, NaN where 0), "
                "R/G/B planes and brick queues resident in HBM",
        "config": {"workload": "loop-closure de-integration of the 640x480, 5 mm walk's frames at corrected "
                               "poses, in the reference's pose order, 50 frames per submission, into a partner "
                               "map sized to the map",
                   "frames": n_frames, "resolution": [w, h], "voxel_m": cfg["scale"],
                   "truncation_m": cfg["threshold"]},
        "bricks_per_s": round(bricks / dt, 1),
        "bricks_per_frame": round(bricks / args.steps, 1),
        "map": {"bricks": counts["n_leaves_used"], "partner_slots_used": lp.mapped()},
        "roofline": {"kernel": "k_loop_delete", "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "avg_launch_us": round(k_ms * 1e3, 3), "bytes_per_launch": round(k_bytes)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_streams(args):
    """Batched streams: every stream advances one frame per tick (one SDF and
    one marching-cubes launch for all of them); aggregate frames/s and the
    tick latency distribution."""
    from fastfusion_amd import Batch, FusionMipMapGPU, hiprt
    n, F = args.streams, args.frames
    from fastfusion_amd import synth
    cfgs = synth.stream_configs(n)
    t_r = time.perf_counter()
    dev, frames_t = [], []
    for k, c in enumerate(cfgs):
        R, t = c["scene"].walk(F, c["seed"] + 7, 0.5, 3.0, 0.02, 2.0)
        w, h = c["w"], c["h"]
        depth = np.zeros((F, h, w), np.uint16)
        bgr = np.zeros((F, h, w, 3), np.uint8)
        for i in range(F):
            depth[i], bgr[i] = c["scene"].render(R[i].reshape(3, 3), t[i], c["K"], w, h, seed=c["seed"] * 10007 + i)
        dev.append((hiprt.DeviceBuffer.from_array(depth), hiprt.DeviceBuffer.from_array(bgr)))
        frames_t.append((R, t))
        del depth, bgr
        print(f"[bench] stream {k}: {w}x{h}, {c['scale'] * 1e3:.1f} mm, room {c['size'][0]:.1f}x{c['size'][1]:.1f}, "
              f"{c['prims']} primitives ({time.perf_counter() - t_r:.0f} s)", file=sys.stderr, flush=True)
    fus = [FusionMipMapGPU(0, 0, 0, c["scale"], c["threshold"], 0, True, max_leaves=1 << 18) for c in cfgs]
    batch = Batch(fus)
    batch.autotune(2)

    def tick(i):
        fr = []
        for k, c in enumerate(cfgs):
            fb, cb = c["w"] * c["h"] * 2, c["w"] * c["h"] * 3
            R, t = frames_t[k]
            fr.append((dev[k][0].addr(i * fb), dev[k][1].addr(i * cb), c["w"], c["h"], R[i], t[i], c["K"],
                       1 / 5000.0, 4.0))
        batch.addMaps(fr)
        batch.updateMeshes()

    warm = min(args.warmup_ticks, F - 1)
    for i in range(warm):
        tick(i)
    hiprt.synchronize()
    cfg_tuned = batch.config()
    print(f"[bench] tuned {cfg_tuned}", file=sys.stderr, flush=True)
    lat = []
    t0 = time.perf_counter()
    for i in range(warm, F):
        a = time.perf_counter()
        tick(i)
        lat.append(time.perf_counter() - a)
    hiprt.synchronize()
    dt = time.perf_counter() - t0
    ticks = F - warm
    # the batched SDF launch against the HBM roofline: a few more ticks with
    # direct launches and HIP events on the lead context (replaying the last frames)
    fus[0].set_timing(True)
    fus[0].reset_timing()
    batch.set_config(cfg_tuned["threads_per_brick"], cfg_tuned["bricks_per_workgroup"], False)
    for i in range(F - min(20, F), F):
        tick(i)
    hiprt.synchronize()
    tm = fus[0].timing()["sdf"]
    sdf_ms = tm["ms"] / max(tm["calls"], 1)
    sdf_bytes = tm["bytes"] / max(tm["calls"], 1)
    achieved = sdf_bytes / (sdf_ms * 1e-3) / 1e9 if sdf_ms > 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    res = {"w": [c["w"] for c in cfgs], "vox_mm": [round(c["scale"] * 1e3, 2) for c in cfgs]}
    line = {
        "metric": "batched streams: aggregate fused+meshed frames/s (24 independent streams, one SDF and one "
                  "marching-cubes launch per tick)",
        "value": round(n * ticks / dt, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": ticks,
        "warmup": warm,
        "ms_per_step": round(dt * 1e3 / ticks, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded analytic rooms ray-cast on the host, resident in HBM",
        "config": {"workload": f"{n} streams x {F} frames, resolutions {{320x240, 640x480, 1280x960}}, "
                               "4-10 mm voxels, 3-voxel truncation, rooms 2-8 m with 3-20 primitives",
                   "streams": res, "launch": cfg_tuned},
        "tick_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 3),
                    "p99": round(float(np.percentile(lat_ms, 99)), 3),
                    "per_stream": "every stream completes at its tick's end: its p50/p99 equal the tick's"},
        "roofline": {"kernel": "k_sdf_batch", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_us": round(sdf_ms * 1e3, 3), "bytes_per_launch": round(sdf_bytes)},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed frames (default 950; 2950 with --large)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed frames before them (default 50)")
    ap.add_argument("--seq-frames", type=int, default=None,
                    help="frames of the sequence (default 1000; 3000 with --large); the timed ones are its last")
    ap.add_argument("--large", action="store_true",
                    help="large-map workload: 1280x960, 2.5 mm floor, learned weighting, 3000 frames")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graphs", action="store_true", help="replay each frame from captured HIP graphs")
    ap.add_argument("--learned", action="store_true",
                    help="learned per-pixel confidence on (committed weights): the throughput-cost run")
    ap.add_argument("--streams", type=int, default=0,
                    help="batched-streams workload with this many streams (24 in the reference workload)")
    ap.add_argument("--frames", type=int, default=300, help="frames per stream (batched-streams workload)")
    ap.add_argument("--warmup-ticks", type=int, default=80, help="untimed ticks (autotuning) of the streams workload")
    ap.add_argument("--loop", action="store_true",
                    help="loop-closure de-integration workload (subtract the walk's frames from a partner map)")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="MiB copied device-to-device once before the run (HBM counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise the counter_collection.csv of a FETCH_SIZE and a WRITE_SIZE rocprofv3 run of "
                         "this script (with --calib-copy 256) for the timed frames given by --steps / --warmup")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 2950 if args.large else 950
    if args.warmup is None:
        args.warmup = 50
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: every run here uses one "
              "GPU (independent map streams would shard by stream, a single map by Morton range of its bricks)",
              file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        from fastfusion_amd import hiprt
        nb = args.calib_copy << 20
        src, dst = hiprt.DeviceBuffer(nb), hiprt.DeviceBuffer(nb)
        hiprt.check(hiprt.hip().hipMemcpy(dst.ptr, src.ptr, nb, hiprt.D2D), "calibration copy")
        hiprt.synchronize()
        del src, dst
    if args.pmc_summary:
        return pmc_summary(args)
    if args.streams:
        return run_streams(args)
    if args.loop:
        return run_loop(args)

    return run_walk(args)


def pmc_summary(args):
    """HBM traffic per frame and per kernel launch of the timed frames, from
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this script (KiB per
    dispatch), each scaled by what the counter reads for the 256 MiB
    device-to-device copy that --calib-copy 256 runs first.  Dispatches are
    grouped into frames by the frame's first kernel (k_planes / k_planes_v:
    a frame's kernels, then its mesh update's, are enqueued in that order);
    the timed frames are those of the timed run (the first pass over the
    walk), i.e. frames [n_walk - steps, n_walk), up to the gather of the
    final mesh that follows them."""
    import csv
    import collections
    import re

    def load(path):
        rows = []
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        rows.sort()
        return rows

    n_walk = max(args.seq_frames or (3000 if args.large else 1000), args.warmup + args.steps)
    first, last = n_walk - args.steps, n_walk
    out = {}
    for tag, path in (("fetch", args.pmc_summary[0]), ("write", args.pmc_summary[1])):
        rows = load(path)
        calib = next(v for _, k, v in rows if "copyBuffer" in k)  # the 256 MiB copy comes first
        scale = 256.0 * 1024.0 / calib  # KiB the copy moved / KiB the counter read
        frame = -1
        per = collections.defaultdict(list)
        total = 0.0
        done = False
        for _, k, v in rows:
            name = k.split("(")[0].replace("void ", "").replace("ff::", "").split("<")[0].strip()
            mg = re.match(r"_ZN2ff(\d+)", name)  # a name the profiler left mangled
            if mg:
                name = name[mg.end():mg.end() + int(mg.group(1))]
            if name in ("k_planes", "k_planes_v"):
                frame += 1
            # the final mesh's gather ends the timed run (the replay follows)
            if frame >= first and name == "k_gather":
                done = True
            if first <= frame < last and not done:
                b = v * 1024.0 * scale
                per[name].append(b)
                total += b
        out[tag] = {"scale": scale, "frame": total / args.steps,
                    "kernels": {k: (sum(v) / len(v), len(v)) for k, v in per.items()}}
    f, w = out["fetch"], out["write"]
    names = sorted(set(f["kernels"]) | set(w["kernels"]), key=lambda k: -(f["kernels"].get(k, (0, 0))[0] * f["kernels"].get(k, (0, 0))[1]))
    print(f"calibration: FETCH_SIZE x{f['scale']:.4f}, WRITE_SIZE x{w['scale']:.4f} (256 MiB device copy)")
    print("| kernel | launches | read B / launch | written B / launch |")
    print("|---|---|---|---|")
    for k in names:
        rb, n = f["kernels"].get(k, (0.0, 0))
        wb, _ = w["kernels"].get(k, (0.0, 0))
        print(f"| {k} | {n} | {rb:,.0f} | {wb:,.0f} |")
    key = "learned" if args.learned or args.large else "plain"
    ks = {k: f["kernels"].get(k, (0.0, 0))[0] + w["kernels"].get(k, (0.0, 0))[0] for k in ("k_sdf", "k_mc")}
    print(f"per frame: read {f['frame']:,.0f} B, written {w['frame']:,.0f} B")
    print(f"bench-traffic: {key} frame={f['frame'] + w['frame']:.0f} k_sdf={ks['k_sdf']:.0f} k_mc={ks['k_mc']:.0f}")


def read_traffic(path, key):
    """HBM traffic figures of a committed counter profile: the line
    `bench-traffic: <key> frame=B k_sdf=B k_mc=B` of the summary (bytes per
    frame and per launch, from calibrated FETCH_SIZE + WRITE_SIZE)."""
    try:
        for line in open(path):
            line = line.strip().strip("`")
            if line.startswith("bench-traffic:"):
                parts = line.split()
                if len(parts) > 1 and parts[1] == key:
                    return {k: float(v) for k, v in (x.split("=") for x in parts[2:])}
    except OSError:
        pass
    return {}


def kernel_entry(name, ms, calls, bytes_per_launch, traffic):
    avg = ms / max(calls, 1)
    ach = bytes_per_launch / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "avg_launch_us": round(avg * 1e3, 3),
            "timed_launches": int(calls), "bytes_per_launch": round(bytes_per_launch)}


def run_walk(args):
    """The 640x480, 5 mm walk (or, with --large, the 1280x960, 2.5 mm floor
    with learned weighting).  Frames before the warm-up and timed ones are
    fused untimed first, so that the timed frames are always the sequence's
    last `steps` frames (its steady state) whatever --steps is.  The
    algorithmic bytes of the timed frames come from a replay of the same
    frames on a second map with every phase evented (the map is
    deterministic: the replay's counts equal the timed run's, which is
    checked).  The frames are rendered on the host in chunks and kept in HBM."""
    from fastfusion_amd import FusionMipMapGPU, hiprt, learned, synth

    large = args.large
    n_seq = max(args.seq_frames or (3000 if large else 1000), args.warmup + args.steps)
    prefuse = n_seq - args.warmup - args.steps
    if large:
        sc, R, t, cfg = synth.config_c5(n_frames=n_seq)
        K = cfg["K"]
        what = "1280x960 2.5 mm floor"
    else:
        sc, R, t, cfg = synth.config_c2(n_frames=n_seq)
        K = synth.K_DEFAULT
        what = "640x480 5 mm"
    w, h = cfg["w"], cfg["h"]
    fb, cb = w * h * 2, w * h * 3
    d_depth = hiprt.DeviceBuffer(fb * n_seq)
    d_bgr = hiprt.DeviceBuffer(cb * n_seq)

    def frame(i):
        return sc.render(R[i].reshape(3, 3), t[i], K, w, h, noise=cfg["noise"], dropout=cfg["dropout"],
                         seed=1000 + i)

    t_r = time.perf_counter()
    chunk = 100
    for c0 in range(0, n_seq, chunk):
        n = min(chunk, n_seq - c0)
        dd = np.zeros((n, h, w), np.uint16)
        bb = np.zeros((n, h, w, 3), np.uint8)
        for k in range(n):
            dd[k], bb[k] = frame(c0 + k)
        hiprt.copy_to_device(d_depth.addr(c0 * fb), dd)
        hiprt.copy_to_device(d_bgr.addr(c0 * cb), bb)
        if large:
            print(f"[bench] rendered {c0 + n}/{n_seq} frames ({time.perf_counter() - t_r:.0f} s)", file=sys.stderr,
                  flush=True)
    print(f"[bench] rendered {n_seq} frames in {time.perf_counter() - t_r:.1f} s", file=sys.stderr, flush=True)
    use_learned = args.learned or large

    def make():
        if large:
            # 8 Mi bricks (57 GiB of TSDF) and 2^27 tree slots: sized for one
            # MI355X's 288 GB beside the frames and the mesh pools
            fz = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 23,
                                 max_branches=1 << 27)
        else:
            fz = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 21,
                                 max_branches=1 << 25)
        if args.graphs:
            fz.setGraphCapture(True)
        if use_learned:
            fz.setLearnedWeighting(learned.load_weights())
        return fz

    def fuse(fz, frames, tag):
        nq = 0
        for k, i in enumerate(frames):
            nq += fz.addMapDevice(d_depth.addr(i * fb), d_bgr.addr(i * cb), w, h, R[i], t[i], K, 1 / 5000.0,
                                  cfg["maxcam"])
            fz.updateMeshes()
            if k % 100 == 99:  # (a profiled run is slow: it shows its progress)
                print(f"[bench] {tag}: {k + 1} frames", file=sys.stderr, flush=True)
        return nq

    untimed = range(0, prefuse + args.warmup)
    timed = range(prefuse + args.warmup, n_seq)
    # the timed run: k_sdf's and k_mc's dispatches record events on one
    # frame in 8 (each event costs launch time)
    fz = make()
    fuse(fz, untimed, "untimed")
    fz.sync()  # (the last requested mesh update enqueued and done)
    fz.reset_timing()
    fz.set_timing(True, sdf_only=True, every=8)
    t0 = time.perf_counter()
    nq_sum = fuse(fz, timed, "timed")
    fz.sync()  # the last frame's mesh update included
    dt = time.perf_counter() - t0
    tm = fz.timing()
    mc = fz.mc_timing()
    counts = fz.counts()
    print(f"[bench] timed {args.steps} frames in {dt:.4f} s", file=sys.stderr, flush=True)
    # surface error of the final mesh against the analytic scene (GT); its
    # gather kernel also ends the timed run in a counter profile
    t1 = time.perf_counter()
    mesh = fz.getMeshInterleavedMarchingCubes()
    t_copy_inc = time.perf_counter() - t1
    err = sc.sdf(mesh.vertices) if len(mesh.vertices) else np.zeros(1, np.float32)
    rms = float(np.sqrt(np.mean(err.astype(np.float64) ** 2)))
    # the final full extraction (getMeshRecursive's counterpart), timed apart
    hiprt.synchronize()
    t1 = time.perf_counter()
    fz.meshFull()
    t_full = time.perf_counter() - t1
    t1 = time.perf_counter()
    full = fz.getMeshInterleavedMarchingCubes()
    t_copy = time.perf_counter() - t1
    same_full = (len(full.faces) == len(mesh.faces) and
                 np.array_equal(full.vertices.view(np.uint32), mesh.vertices.view(np.uint32)))
    del full
    n_cells = len(fz.getMeshCells()[0])
    del fz

    # the replay: every phase evented over the same timed frames
    fr = make()
    fuse(fr, untimed, "replay untimed")
    fr.sync()
    fr.reset_timing()  # the counts cover the timed frames only
    fr.set_timing(True)
    nq_replay = fuse(fr, timed, "replay")
    fr.sync()
    rp = fr.timing()
    rmc = fr.mc_timing()
    rc = fr.counts()
    same = nq_replay == nq_sum and all(rc[k] == counts[k] for k in ("n_leaves_used", "n_branches_used", "grows"))
    if not same:
        raise SystemExit("[bench] the replay's map differs from the timed run's")
    del fr

    # whole frame (addMap + updateMeshes): algorithmic bytes of the timed frames
    B = sum(v["bytes"] for v in rp.values())
    achieved = B / dt / 1e9
    tr = read_traffic(os.path.join(ROOT, "profiles", "hbm_traffic_1280x960_2.5mm_floor.md" if large else
                                   "hbm_traffic_640x480_5mm.md"), "learned" if use_learned else "plain")
    sdf = tm["sdf"]
    sdf_bpl = sdf["bytes"] / max(sdf["calls"], 1)
    mc_bpl = rmc["bytes"] / max(rmc["byte_calls"], 1)
    kernels = {
        "k_sdf": kernel_entry("k_sdf", sdf["ms"], sdf["calls"], sdf_bpl, tr.get("k_sdf")),
        "k_mc": kernel_entry("k_mc", mc["ms"], mc["calls"], mc_bpl, tr.get("k_mc")),
    }
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(frame, n_seq, R, t, K, cfg, args.cpu_seconds, what)
    nst = max(args.steps, 1)
    if large:
        metric = ("large map: fused+meshed frames/sec and HBM GB/s at 1280x960, 2.5mm, learned weighting; "
                  "RMS surface err vs GT; final full extraction time")
        workload = (f"large map: 12x12x3 m floor of 3x3 rooms, 60 primitives (seed 3000), {n_seq}-frame seeded "
                    "random walk, 1280x960 (f = 1050), 2.5 mm voxels, 10 mm truncation, 4 m max depth, learned "
                    "confidence on; timed: its last frames")
    else:
        metric = METRIC
        workload = (("learned confidence" if args.learned else "plain method") +
                    f", 640x480, 5 mm voxels, 15 mm truncation, 4 m max depth, {n_seq}-frame seeded random walk "
                    "(room seed 11, walk seed 12); timed: its last frames")
    line = {
        "metric": metric,
        "value": round(args.steps / dt, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / nst, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded analytic scene ray-cast on the host, u16 depth + BGR8, resident in HBM",
        "config": {"workload": workload,
                   "frames": n_seq, "untimed_before_warmup": prefuse, "timed_frames": [timed.start, timed.stop],
                   "resolution": [w, h], "voxel_m": cfg["scale"], "truncation_m": cfg["threshold"],
                   "maxcam_m": cfg["maxcam"], "learned": use_learned},
        "roofline": {"kernel": "whole frame (addMap + updateMeshes, every kernel)", "bound": "hbm",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr.get("frame"),
                     "algorithmic_bytes_per_frame": round(B / nst),
                     "bytes_split_per_frame": {"img_alloc": round(rp["alloc"]["bytes"] / nst),
                                               "sdf": round(rp["sdf"]["bytes"] / nst),
                                               "mesh": round(rp["mesh"]["bytes"] / nst)},
                     "kernels": kernels},
        "bricks_per_frame": round(nq_sum / nst, 1),
        "processed_px_per_frame": round(rp["alloc"]["units"] / nst, 1),
        "remeshed_cells_per_frame": round(rp["mesh"]["units"] / nst, 1),
        "map": {"bricks": counts["n_leaves_used"], "side": counts["side"], "cells": n_cells,
                "tree_slots": counts["n_branches_used"]},
        "phases_ms_per_frame_replay": {k: round(v["ms"] / max(v["calls"], 1), 4) for k, v in rp.items()},
        "full_extraction": {"ms": round(t_full * 1e3, 2), "copy_to_host_ms": round(t_copy * 1e3, 2),
                            "equals_incremental": bool(same_full), "incremental_copy_ms": round(t_copy_inc * 1e3, 2)},
        "rms_surface_err_m": round(rms, 6),
        "rms_surface_err_voxels": round(rms / cfg["scale"], 4),
        "mesh": {"vertices": int(len(mesh.vertices)), "triangles": int(len(mesh.faces))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
