"""GPU parity of incremental meshing: after each frame, the GPU's per-cell
meshes (kept across frames, only dirty cells re-meshed) must equal the
oracle's meshes of every cell computed from scratch: same cell set
(type, origin, size), same vertex and face counts per cell, identical faces
and colors, vertices within 1e-4 voxel."""
import numpy as np
import pytest

from fastfusion_amd import synth
from tests import helpers
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu


def compare_mesh(gpu, ora, voxel, tag):
    d_g, nv_g, nf_g = gpu.getMeshCells()
    desc, nv_o, nf_o, xyz_o, rgb_o, faces_o = ora.mesh_all()
    assert len(d_g) == len(desc), f"{tag}: {len(d_g)} gpu cells vs {len(desc)} oracle cells"
    assert np.array_equal(d_g, desc[:, :5]), f"{tag}: cell descriptors differ"
    bad = np.flatnonzero((nv_g != nv_o) | (nf_g != nf_o))
    assert bad.size == 0, (f"{tag}: {bad.size} cells differ in counts, first {desc[bad[:3], :5]} "
                           f"gpu {nv_g[bad[:3]]}/{nf_g[bad[:3]]} oracle {nv_o[bad[:3]]}/{nf_o[bad[:3]]}")
    m = gpu.getMeshInterleavedMarchingCubes()
    assert len(m.vertices) == len(xyz_o) and len(m.faces) == len(faces_o)
    err = np.abs(m.vertices.astype(np.float64) - xyz_o).max() if len(xyz_o) else 0.0
    assert err <= 1e-4 * voxel, f"{tag}: vertex error {err}"
    # oracle faces are local per cell; make them global
    off = np.repeat(np.concatenate([[0], np.cumsum(nv_o)[:-1]]).astype(np.uint32), nf_o)
    assert np.array_equal(m.faces, faces_o + off[:, None]), f"{tag}: faces differ"
    assert np.array_equal(m.colors, rgb_o), f"{tag}: colors differ"
    return len(desc), len(m.faces)


def run_mesh(sc, R, t, frames, cfg, K, check_every=1, max_leaves=1 << 17, sample_tiles=None, pool=None, out_gpu=None):
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=max_leaves)
    if sample_tiles is not None:
        gpu.setMeshSampleTiles(sample_tiles)
    if pool is not None:
        gpu.setMeshPool(*pool)
    if out_gpu is not None:
        out_gpu.append(gpu)
    w, h = cfg["w"], cfg["h"]
    out = None
    for k, i in enumerate(frames):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, w, h, noise=cfg.get("noise", 1.0),
                           dropout=cfg.get("dropout", 0.005), seed=1000 + i)
        assert gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"]) == \
            ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        gpu.updateMeshes()
        if k % check_every == 0 or k == len(frames) - 1:
            out = compare_mesh(gpu, ora, cfg["scale"], f"frame {i}")
    return out


def test_mesh_c1_mixed_scales():
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    cfg = dict(cfg, w=320, h=240)
    ncell, nface = run_mesh(sc, R, t, list(range(0, 12)), cfg, (262.5, 262.5, 159.5, 119.5), check_every=3)
    assert nface > 1000


def test_mesh_c1_640():
    """Configuration 1 at its own 640x480 resolution (the mixed-scale test
    above runs it at 320x240): meshes equal the oracle's every third frame."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    ncell, nface = run_mesh(sc, R, t, list(range(0, 40, 4)), cfg, synth.K_DEFAULT, check_every=3)
    assert nface > 10000


def test_mesh_c2_walk_5mm():
    sc, R, t, cfg = synth.config_c2(n_frames=60)
    ncell, nface = run_mesh(sc, R, t, list(range(0, 16)), cfg, synth.K_DEFAULT, check_every=5, max_leaves=1 << 18)
    assert nface > 10000


def test_mesh_close_uniform_scale():
    sc = synth.Scene.room_scene(5, (4.0, 4.0, 3.0), 3, 3, 0.1, 0.4)
    f = np.array([-1.0, 0.15, -0.1]); f /= np.linalg.norm(f)
    x = np.cross(f, [0, 0, 1.0]); x /= np.linalg.norm(x)
    y = np.cross(f, x)
    Rm = np.stack([x, y, f], 1).ravel()
    R = np.tile(Rm, (6, 1))
    t = np.array([[0.8, 2.0 + 0.01 * i, 1.5] for i in range(6)])
    cfg = dict(scale=0.01, threshold=0.03, maxcam=1.1, w=320, h=240)
    run_mesh(sc, R, t, list(range(6)), cfg, (262.5, 262.5, 159.5, 119.5), check_every=1)


def test_mesh_after_grow():
    from tests.test_cpu_oracle import grow_sequence
    sc, R, t = grow_sequence()
    cfg = dict(scale=0.02, threshold=0.06, maxcam=6.0, w=320, h=240)
    run_mesh(sc, R, t, list(range(len(R))), cfg, (262.5, 262.5, 159.5, 119.5), check_every=1)


def test_mesh_sample_buffer_overflow():
    """A sample buffer of 8 tiles cannot hold the first update's cells: the
    cells beyond it are meshed in place by the second pass, the buffer grows,
    and every frame's meshes still equal the oracle's."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    cfg = dict(cfg, w=320, h=240)
    ncell, nface = run_mesh(sc, R, t, list(range(0, 6)), cfg, (262.5, 262.5, 159.5, 119.5), check_every=1,
                            sample_tiles=8)
    assert nface > 1000


def test_mesh_sample_buffer_overflow_beside_next_frame():
    """The same 8-tile buffer with the meshes read only every third frame:
    the first update is enqueued by the next frame, whose SDF update follows
    that update's sampling and must stand down for the host (the cells that
    did not fit are sampled from the bricks by the marching cubes, which may
    still run); the meshes read later still equal the oracle's."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    gpu.setMeshSampleTiles(8)
    for k in range(7):
        d, bgr = sc.render(R[k].reshape(3, 3), t[k], K, 320, 240, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + k)
        assert gpu.addMap(d, R[k], t[k], K, bgr, 1 / 5000.0, cfg["maxcam"]) == \
            ora.add_map(d, R[k], t[k], K, bgr, 1 / 5000.0, cfg["maxcam"])
        gpu.updateMeshes()
        if k in (3, 6):
            ncell, nface = compare_mesh(gpu, ora, cfg["scale"], f"frame {k}")
    assert nface > 1000


def test_mesh_unchanged_tree_remarks_in_place():
    """Frames that create no brick or branch leave the cell partition as it
    was: the update then re-marks the previous table in place.  Re-fusing
    frames already seen guarantees such updates; the meshes must still equal
    the oracle's from-scratch meshes, including cells whose new mesh is
    empty."""
    sc, R, t, cfg = synth.config_c2(n_frames=60)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    frames = list(range(0, 6)) + [5, 5, 3, 4, 5, 2]
    unchanged = 0
    prev = None
    for k, i in enumerate(frames):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + i + (100 * k if k >= 6 else 0))
        assert gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"]) == \
            ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        c = gpu.counts()
        key = (c["n_leaves_used"], c["n_branches_used"], c["grows"])
        unchanged += key == prev
        prev = key
        gpu.updateMeshes()
        compare_mesh(gpu, ora, cfg["scale"], f"step {k} (frame {i})")
    assert unchanged >= 3


def test_mesh_compaction_behind_overlapped_frames():
    """Updates that compact the live geometry into fresh pools (arenas ask
    for a compaction once 1 % of them is used), compared only after the last
    frame: every update but the last then finishes (its host part, with the
    compaction) after the next frame took its counters up, without waiting
    for the update's own stream.  The meshes must still equal the oracle's."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    frames = list(range(0, 24, 2))
    got = []
    run_mesh(sc, R, t, frames, cfg, K, check_every=len(frames), pool=(1 << 18, 1), out_gpu=got)
    cap, compactions = got[0].meshPoolInfo()
    print(f"mesh compaction: pools of {cap}, {compactions} compactions", flush=True)
    assert cap >= 1 << 18 and compactions >= 1  # (an update that overflowed doubles the pools)


def test_mesh_update_beside_next_frame():
    """Each update is enqueued and the next frame fused at once: its
    allocation, which links new branches and bricks into the tree and stamps
    the bricks it queues, runs beside the update's partition and marking.
    The update must mesh the map as it was when it was enqueued (links to
    nodes past the counts it was launched with read as empty), so its mesh,
    read after the next frame returned, equals the oracle's mesh of the map
    before that frame."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    frames = list(range(0, 48, 3))
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    w, h = cfg["w"], cfg["h"]

    def frame(i):
        return sc.render(R[i].reshape(3, 3), t[i], K, w, h, noise=cfg["noise"], dropout=cfg["dropout"],
                         seed=1000 + i)

    d, bgr = frame(frames[0])
    gpu.addMap(d, R[frames[0]], t[frames[0]], K, bgr, 1 / 5000.0, cfg["maxcam"])
    ora.add_map(d, R[frames[0]], t[frames[0]], K, bgr, 1 / 5000.0, cfg["maxcam"])
    grew = 0
    for a, b in zip(frames[:-1], frames[1:]):
        before = gpu.counts()["n_branches_used"]
        gpu.updateMeshes()
        d, bgr = frame(b)
        nq = gpu.addMap(d, R[b], t[b], K, bgr, 1 / 5000.0, cfg["maxcam"])
        grew += gpu.counts()["n_branches_used"] > before
        compare_mesh(gpu, ora, cfg["scale"], f"update of frame {a}, frame {b} beside it")
        assert nq == ora.add_map(d, R[b], t[b], K, bgr, 1 / 5000.0, cfg["maxcam"])
    assert grew >= 5  # the frames beside the updates linked new branches
    gpu.updateMeshes()
    compare_mesh(gpu, ora, cfg["scale"], f"frame {frames[-1]}")


def test_mesh_pools_too_small_grow():
    """Pools of 4 096 vertices, far below the live mesh: updates overflow,
    their compactions lose geometry, so the pools double and the next update
    meshes every cell again; once an update fits, the meshes equal the
    oracle's."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 17)
    gpu.setMeshPool(4096, 50)
    for i in range(0, 12, 2):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, cfg["w"], cfg["h"], noise=1.0, dropout=0.005, seed=1000 + i)
        assert gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"]) == \
            ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        gpu.updateMeshes()
    caps = [gpu.meshPoolInfo()[0]]
    for _ in range(10):  # no new frame: an update after an overflow meshes every cell
        gpu.updateMeshes()
        caps.append(gpu.meshPoolInfo()[0])
        if len(caps) >= 3 and caps[-1] == caps[-2] == caps[-3]:
            break
    print(f"mesh pools: {caps}", flush=True)
    assert caps[-1] > 4096 and caps[-1] == caps[-2]
    compare_mesh(gpu, ora, cfg["scale"], "after growth")


def test_cell_update_matches_oracle_incremental():
    """getMeshCellsUpdate (geometryfusion_mipmap_cpu.cpp:2989-3047): after
    each frame and updateMeshes, the cells the GPU re-meshed are the cells
    the oracle's incremental mesher re-meshed, in key order, each with the
    same mesh (faces local to the cell)."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    sizes = []
    for k, i in enumerate([0, 3, 6, 9, 9, 12, 40]):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + i + 100 * k)
        assert gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"]) == \
            ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        gpu.updateMeshes()
        ora.update_meshes()
        u = gpu.getMeshCellsUpdate()
        desc, nv, nf, xyz, rgb, faces = ora.cell_update()
        tag = f"update {k} (frame {i})"
        assert np.array_equal(u.desc, desc[:, :5]), f"{tag}: {len(u.desc)} gpu cells vs {len(desc)} oracle cells"
        assert np.array_equal(u.n_vertices, nv) and np.array_equal(u.n_faces, nf), tag
        err = np.abs(u.vertices.astype(np.float64) - xyz).max() if len(xyz) else 0.0
        assert err <= 1e-4 * cfg["scale"], f"{tag}: vertex error {err}"
        assert np.array_equal(u.faces, faces) and np.array_equal(u.colors, rgb), tag
        for dsc, v, c, f in u.cells():
            assert len(f) == 0 or f.max() < len(v)
        sizes.append(len(u.desc))
    assert min(sizes) > 0 and max(sizes) > 100
    # fewer cells than the map has: the update is incremental
    assert sizes[-2] < len(gpu.getMeshCells()[0])


def test_full_mesh_equals_incremental_union():
    """The from-scratch mesh of the whole map (ff_mesh_full, the counterpart
    of getMeshRecursive) equals the union of the incremental cell meshes,
    bit for bit, and the oracle's from-scratch meshes."""
    sc, R, t, cfg = synth.config_c2(n_frames=60)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    for i in range(0, 24, 2):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + i)
        gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        gpu.updateMeshes()
    inc = gpu.getMeshInterleavedMarchingCubes()
    gpu.meshFull()
    full = gpu.getMeshInterleavedMarchingCubes()
    assert len(full.faces) > 1000
    assert np.array_equal(inc.vertices.view(np.uint32), full.vertices.view(np.uint32))
    assert np.array_equal(inc.colors, full.colors) and np.array_equal(inc.faces, full.faces)
    # every cell was re-meshed by the full pass
    assert len(gpu.getMeshCellsUpdate().desc) == len(gpu.getMeshCells()[0])
    compare_mesh(gpu, ora, cfg["scale"], "full")
