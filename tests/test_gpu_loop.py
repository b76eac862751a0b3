"""GPU parity of the loop-closure de-integration (the ff_loop_* entry points:
updateWrapper_deleteloop + update8AddLoopSSESingleDeleteloop,
loopclosure.cpp:15-311; deleteOldLoopPoses,
geometryfusion_mipmap_cpu.cpp:3934-4024) against the oracle, bit for bit:
slot hand-out (_leafMapLoop, _nLeavesMapped), the partner's distances,
weights and colors, the stop when the slots run out, duplicates in a queue,
maps without color, the device-pointer entry point and the reset."""
import numpy as np
import pytest

from fastfusion_amd import FFError, hiprt, loopclosure, synth
from fastfusion_amd.loopclosure import LoopClosure
from oracle.pyoracle import LoopOracle
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu

K = synth.K_DEFAULT


def _fused(n, color=True, w=640, h=480, K_=K):
    """n frames fused into a GPU map and an oracle map; per frame the raw
    images, the pose and the queue recorded after it (_leafQueues)."""
    sc, R, t, cfg = synth.config_c1(n_frames=60, w=w, h=h)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], color=color)
    frames = []
    for i in range(n):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K_, w, h, seed=500 + i)
        ng = gpu.addMap(d, R[i], t[i], K_, bgr if color else None, 1 / 5000.0, cfg["maxcam"])
        no = ora.add_map(d, R[i], t[i], K_, bgr if color else None, 1 / 5000.0, cfg["maxcam"])
        assert ng == no
        q = gpu.queue()
        assert np.array_equal(q, np.sort(ora.queue()))
        frames.append((d, bgr, R[i].reshape(3, 3), t[i], q))
    return gpu, ora, frames


def _moved(R, t, k):
    """A loop-closure correction of pose k: 1-3 cm and under a degree."""
    a = 0.004 * (k + 1)
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    return Rz @ R, np.asarray(t, np.float64) + np.array([0.01, -0.005 * k, 0.02])


def _same(lg, lo, tag, n_leaves):
    assert lg.mapped() == lo.mapped(), f"{tag}: mapped {lg.mapped()} vs {lo.mapped()}"
    mg, mo = lg.slot_map(n_leaves), lo.slot_map()[:n_leaves]
    assert np.array_equal(mg, mo), f"{tag}: slot maps differ at {np.flatnonzero(mg != mo)[:6]}"
    slots = np.arange(lo.mapped(), dtype=np.uint32)
    dg, wg, cg = lg.bricks(slots)
    do, wo, co = lo.bricks(slots)
    bad = np.flatnonzero(wg.view(np.uint32).ravel() != wo.view(np.uint32).ravel())
    assert bad.size == 0, f"{tag}: weights differ in {bad.size} voxels, first {bad[:4]}"
    bad = np.flatnonzero(dg.view(np.uint32).ravel() != do.view(np.uint32).ravel())
    assert bad.size == 0, f"{tag}: distances differ in {bad.size} voxels: {dg.ravel()[bad[:3]]} vs {do.ravel()[bad[:3]]}"
    if lg.color:
        assert np.array_equal(cg, co), f"{tag}: colors differ in {(cg != co).sum()} values"
    return wg


def test_delete_frames_640_matches_oracle():
    print("loop: 640x480, 6 frames, color", flush=True)
    gpu, ora, frames = _fused(6)
    n = gpu.counts()["n_leaves_used"]
    lg, lo = LoopClosure(gpu), LoopOracle(ora, loopclosure.MAX_LOOP_CLOSURE_LEAVES)
    for k, (d, bgr, R, t, q) in enumerate(frames):
        R2, t2 = _moved(R, t, k)
        dm = loopclosure.depth_metres(d)
        assert lg.delete_frame(dm, bgr, R2, t2, K, q) == (True, len(q))
        assert lo.delete_frame(dm, loopclosure.planes(bgr), R2, t2, K, q) == (True, len(q))
        w = _same(lg, lo, f"frame {k}", n)
    # the subtraction's signature: negative fall-off weights, front weights
    # counted up (the reference adds +1 in front of the surface)
    assert (w < 0).sum() > 1000 and (w >= 1).sum() > 1000
    lg.reset()
    assert lg.mapped() == 0 and np.all(lg.slot_map(n) == loopclosure.BRANCHINIT)
    assert not np.any(lg.bricks(np.arange(8, dtype=np.uint32))[1])


def test_order_duplicates_and_exhaustion():
    print("loop: 320x240 permuted queues, duplicates, slots for 2/3 of the bricks", flush=True)
    K2 = (262.5, 262.5, 159.5, 119.5)
    gpu, ora, frames = _fused(4, w=320, h=240, K_=K2)
    n = gpu.counts()["n_leaves_used"]
    union = np.unique(np.concatenate([f[4] for f in frames]))
    total = len(union) * 2 // 3  # the slots run out during the last frames
    assert total > 100
    lg, lo = LoopClosure(gpu, total), LoopOracle(ora, total)
    rng = np.random.default_rng(9)
    filled = False
    for k, (d, bgr, R, t, q) in enumerate(frames):
        q = q[rng.permutation(len(q))]
        q = np.concatenate([q, q[: len(q) // 7], q[:3]]).astype(np.uint32)  # later entries repeat bricks
        dm = loopclosure.depth_metres(d)
        rg = lg.delete_frame(dm, bgr, R, t, K2, q)
        ro = lo.delete_frame(dm, loopclosure.planes(bgr), R, t, K2, q)
        assert rg == ro, f"frame {k}: gpu {rg} oracle {ro}"
        filled |= not rg[0]
        _same(lg, lo, f"frame {k}", n)
    assert filled and lg.mapped() == total


def test_no_color_and_device_entry_point():
    print("loop: 320x240 without color, device pointers", flush=True)
    K2 = (262.5, 262.5, 159.5, 119.5)
    gpu, ora, frames = _fused(3, color=False, w=320, h=240, K_=K2)
    n = gpu.counts()["n_leaves_used"]
    lg, lo, ld = LoopClosure(gpu, 20000), LoopOracle(ora, 20000), LoopClosure(gpu, 20000)
    for k, (d, _, R, t, q) in enumerate(frames):
        R2, t2 = _moved(R, t, k)
        dm = loopclosure.depth_metres(d)
        assert lg.delete_frame(dm, None, R2, t2, K2, q)[0]
        assert lo.delete_frame(dm, None, R2, t2, K2, q)[0]
        dd, qd = hiprt.DeviceBuffer.from_array(dm), hiprt.DeviceBuffer.from_array(q)
        assert ld.delete_frame_device(dd.ptr.value, 0, 0, 0, 320, 240, R2, t2, K2, qd.ptr.value, len(q)) == \
            (True, len(q))
        _same(lg, lo, f"frame {k}", n)
        _same(ld, lo, f"device frame {k}", n)


def test_invalid_brick_rejects_the_call():
    print("loop: invalid brick index", flush=True)
    K2 = (262.5, 262.5, 159.5, 119.5)
    gpu, ora, frames = _fused(2, w=320, h=240, K_=K2)
    n = gpu.counts()["n_leaves_used"]
    d, bgr, R, t, q = frames[0]
    lg = LoopClosure(gpu, 10000)
    dm = loopclosure.depth_metres(d)
    lg.delete_frame(dm, bgr, R, t, K2, q[:50])
    before = lg.bricks(np.arange(50, dtype=np.uint32))
    bad = np.concatenate([q[50:60], [n + 5], q[:10]]).astype(np.uint32)
    with pytest.raises(FFError, match="beyond the map"):
        lg.delete_frame(dm, bgr, R, t, K2, bad)
    assert lg.mapped() == 50
    after = lg.bricks(np.arange(50, dtype=np.uint32))
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    # and the partner still works afterwards
    assert lg.delete_frame(dm, bgr, R, t, K2, q[50:60]) == (True, 10) and lg.mapped() == 60


def test_delete_old_poses_follows_the_pose_order():
    print("loop: deleteOldLoopPoses over 5 poses, 3 visited", flush=True)
    K2 = (262.5, 262.5, 159.5, 119.5)
    gpu, ora, frames = _fused(5, w=320, h=240, K_=K2)
    n = gpu.counts()["n_leaves_used"]
    fused = [(R, t) for (_, _, R, t, _) in frames]
    new = [_moved(R, t, k) if k % 2 == 0 else (R, t) for k, (R, t) in enumerate(fused)]
    queues = [f[4] for f in frames]
    lg = LoopClosure(gpu, 30000)
    posesFused = list(fused)
    order = lg.delete_old_poses(lambda p: (frames[p][0], frames[p][1]), new, posesFused, queues, K2, max_poses=3)
    assert order == loopclosure.pose_order([p[1] for p in new], [p[1] for p in fused], 3)
    assert order[:3] == [4, 2, 0]  # largest translation change first
    for p in order:
        assert posesFused[p] is new[p]
    lo = LoopOracle(ora, 30000)
    for p in order:
        d, bgr, _, _, q = frames[p]
        lo.delete_frame(loopclosure.depth_metres(d), loopclosure.planes(bgr), new[p][0], new[p][1], K2, q)
    _same(lg, lo, "deleteOldLoopPoses", n)


def test_one_submission_equals_frame_by_frame():
    print("loop: 6 frames in one submission vs one by one, a rejected frame, slots running out", flush=True)
    K2 = (262.5, 262.5, 159.5, 119.5)
    gpu, ora, frames = _fused(6, w=320, h=240, K_=K2)
    n = gpu.counts()["n_leaves_used"]
    union = np.unique(np.concatenate([f[4] for f in frames]))
    total = len(union) * 3 // 4
    la, lb, lo = LoopClosure(gpu, total), LoopClosure(gpu, total), LoopOracle(ora, total)
    keep, sub = [], []
    for k, (d, bgr, R, t, q) in enumerate(frames):
        R2, t2 = _moved(R, t, k)
        dm = loopclosure.depth_metres(d)
        pl = loopclosure.planes(bgr)
        if k == 2:  # a brick beyond the map: this frame is skipped, the others run
            q = np.concatenate([q[:5], [n + 3], q[5:]]).astype(np.uint32)
        else:
            lo.delete_frame(dm, pl, R2, t2, K2, q)
        bufs = [hiprt.DeviceBuffer.from_array(a) for a in (dm, pl[0], pl[1], pl[2], q)]
        keep.append(bufs)
        sub.append((bufs[0].addr(), bufs[1].addr(), bufs[2].addr(), bufs[3].addr(), 320, 240, R2, t2, K2,
                    bufs[4].addr(), len(q)))
    with pytest.raises(FFError, match="frame 2"):
        la.delete_frames(sub)
    stopped = False
    for k, f in enumerate(sub):
        if k == 2:
            with pytest.raises(FFError, match="beyond the map"):
                lb.delete_frame_device(*f)
            continue
        stopped |= not lb.delete_frame_device(*f)[0]
    assert stopped
    _same(la, lo, "one submission", n)
    _same(lb, lo, "frame by frame", n)
