"""CPU checks that pin the oracle (the CPU restatement of the reference path)
without running the reference: closed forms, brute force in double
precision, fixed-point equations of the de-dup recurrence, exhaustive
agreement of the RCPPS table with the host instruction, structural
invariants of the octree, and global marching cubes on uniform-scale maps.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from fastfusion_amd import synth
from oracle import pyoracle
from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = (262.5, 262.5, 159.5, 119.5)


# ---------------------------------------------------------------- RCPPS
def test_rcp_table_error_bound():
    # Intel documents |rel err| <= 1.5 * 2^-12 for RCPPS
    x = np.float32(1.0) + np.arange(2048, dtype=np.float32) * np.float32(2.0 ** -11)
    r = pyoracle.rcp(x)
    rel = np.abs(r.astype(np.float64) * x.astype(np.float64) - 1.0)
    assert rel.max() <= 1.5 * 2.0 ** -12
    # exponent shift and sign: rcp(-2^k m) = -2^-k rcp(m)
    rng = np.random.default_rng(1)
    m = (1.0 + rng.random(200)).astype(np.float32)
    for k in (-20, -3, 0, 5, 30):
        a = pyoracle.rcp(m * np.float32(2.0 ** k))
        b = pyoracle.rcp(m) * np.float32(2.0 ** -k)
        assert np.array_equal(a, b)
        assert np.array_equal(pyoracle.rcp(-m * np.float32(2.0 ** k)), -a)
    assert np.isinf(pyoracle.rcp(np.float32([0.0]))[0])
    assert pyoracle.rcp(np.float32([np.inf]))[0] == 0.0


def test_rcp_table_matches_host_instruction():
    """The generator compares the table-driven emulation with the host's
    RCPPS over all 2^32 inputs; a host whose RCPPS differs from the capture
    host is reported as a skip, not a failure."""
    src = os.path.join(ROOT, "fastfusion_amd", "csrc", "gen_rcp_table.c")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        subprocess.check_call(["gcc", "-O2", "-mavx", src, "-o", exe])
        out = os.path.join(d, "t.h")
        r = subprocess.run([exe, out], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            pytest.skip("host RCPPS differs from the captured table: " + r.stderr.strip())
        committed = open(os.path.join(ROOT, "fastfusion_amd", "csrc", "rcp_table.h")).read()
        assert open(out).read() == committed


# ---------------------------------------------------------------- closed forms
def test_leaf_scale_func_closed_form():
    # SCALE_LINEAR: floor(max(z, 1) * 8), so the brick scale steps at 1.125,
    # 2.125 and 4.125 m (next power of two of target / 8)
    for z, want in [(0.3, 8), (1.0, 8), (1.124, 8), (1.125, 9), (1.99, 15), (2.0, 16), (3.0, 24)]:
        assert pyoracle.leaf_scale_func(z) == want
    zs = np.linspace(0.1, 6.0, 2001, dtype=np.float32)
    got = np.array([pyoracle.leaf_scale_func(float(z)) for z in zs])
    want = np.floor(np.maximum(zs, np.float32(1.0)) * np.float32(8.0)).astype(int)
    assert np.array_equal(got, want)


def _plane_image(z, w=64, h=48):
    return np.full((h, w), int(round(z * 5000)), np.uint16)


def test_sdf_update_plane_closed_form():
    """Identity pose, fronto-parallel plane at depth Z: for a voxel at camera
    position p the reference's ray SDF is |p| (1 - Z / p_z); weight 1 in front
    (d < 5 mm), linear fall-off behind, 0 beyond 3 tau."""
    Z, scale, tau = 1.0, 0.01, 0.03
    depth = _plane_image(Z)
    bgr = np.full((48, 64, 3), (10, 20, 30), np.uint8)
    Kc = np.float32([60.0, 60.0, 31.5, 23.5])
    m = np.float32([1, 0, 0, -0.04, 0, 1, 0, -0.04, 0, 0, 1, 0.0])  # camera <- world (voxel units scaled)
    pos = np.uint16([0, 0, 92])  # z from 0.92 to 0.99 m
    dist = np.zeros(512, np.float32); wgt = np.zeros(512, np.float32); col = np.zeros(512 * 3, np.uint16)
    pyoracle.update_brick(depth, bgr, 1 / 5000.0, 4.0, m, Kc, scale, tau, pos, 1, dist, wgt, col)
    idx = np.arange(512)
    px = -0.04 + (idx & 7) * scale; py = -0.04 + ((idx >> 3) & 7) * scale; pz = (92 + (idx >> 6)) * scale
    L = np.sqrt(px * px + py * py + pz * pz)
    dI = L * (1.0 - Z / pz)
    # dI = len - (len * rcp(pz)) * h cancels two ~1 m terms: the RCPPS relative
    # error (<= 3.7e-4) becomes an absolute error of about 3.7e-4 * len / pz * h
    tol = 3.7e-4 * L / pz * Z + 1e-6
    assert np.all(np.abs(dist - np.clip(dI, -tau, tau)) <= tol)
    w_ref = np.where(dI < 0.005, 1.0, np.where(dI < 3 * tau, (3 * tau - dI) / (3 * tau - 0.005), 0.0))
    # away from the 5 mm / 3 tau breakpoints the weight follows from dI
    away = (np.abs(dI - 0.005) > 2 * tol) & (np.abs(dI - 3 * tau) > 2 * tol)
    assert np.all(np.abs(wgt - w_ref)[away] <= 2 * tol[away] / (3 * tau - 0.005) + 1e-3)
    assert np.all(wgt[dI < -0.001] == 1.0)
    # color channels: R from byte 2, G byte 1, B byte 0, times 256
    # c = inc * w * rcp(w) truncated, and RCPPS(1.0) = 0.99976: slightly below 256 B
    c = col.reshape(512, 3)[wgt > 0].astype(int)
    ref = np.array([30 * 256, 20 * 256, 10 * 256])
    assert np.all((c <= ref) & (c >= ref - 4))


def test_sdf_update_brute_force_double():
    """Random pose and depth: the oracle's round-toward-zero + RCPPS update
    agrees with a double-precision evaluation of the same projective SDF to
    the RCPPS error (3e-4 relative), voxel by voxel, over three frames."""
    rng = np.random.default_rng(7)
    w, h = 80, 60
    Kc = np.float32([70.0, 70.0, 39.5, 29.5])
    scale, tau = 0.01, 0.03
    depth = (5000 * (1.1 + 0.05 * rng.standard_normal((h, w)))).astype(np.uint16)
    depth[rng.random((h, w)) < 0.05] = 0
    bgr = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    a = 0.07
    Rm = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]], np.float32)
    t = np.float32([-0.03, -0.02, 0.01])
    m = np.concatenate([Rm, t[:, None]], 1).astype(np.float32).ravel()
    pos = np.uint16([0, 0, 100])
    dist = np.zeros(512, np.float32); wgt = np.zeros(512, np.float32); col = np.zeros(512 * 3, np.uint16)
    D = np.zeros(512); Wt = np.zeros(512)
    idx = np.arange(512)
    g = np.stack([idx & 7, (idx >> 3) & 7, idx >> 6], 1).astype(np.float64) + pos.astype(np.float64)
    P = (g * scale) @ Rm.astype(np.float64).T + t.astype(np.float64)
    L = np.linalg.norm(P, axis=1)
    imx = np.trunc(Kc[2] + Kc[0] * P[:, 0] / P[:, 2]).astype(int)
    imy = np.trunc(Kc[3] + Kc[1] * P[:, 1] / P[:, 2]).astype(int)
    for f in range(3):
        pyoracle.update_brick(depth, bgr, 1 / 5000.0, 4.0, m, Kc, scale, tau, pos, 1, dist, wgt, col)
        ok = (P[:, 2] > 0) & (imx >= 0) & (imx < w) & (imy >= 0) & (imy < h)
        hh = depth[np.clip(imy, 0, h - 1), np.clip(imx, 0, w - 1)] / 5000.0
        dI = L - (L / P[:, 2]) * hh
        wI = np.where(dI < 0.005, 1.0, np.where(dI < 3 * tau, (3 * tau - dI) / (3 * tau - 0.005), 0.0)) * ok
        Wn = Wt + wI
        D = np.where(Wn > 0, (D * Wt + np.clip(dI, -tau, tau) * wI) / np.where(Wn > 0, Wn, 1), 0.0)
        Wt = Wn
        # pixel index from truncation can differ only where the projection is
        # within rounding of a pixel boundary: exclude those voxels
        fx = Kc[2] + Kc[0] * P[:, 0] / P[:, 2]; fy = Kc[3] + Kc[1] * P[:, 1] / P[:, 2]
        stable = (np.abs(fx - np.round(fx)) > 1e-3) & (np.abs(fy - np.round(fy)) > 1e-3)
        # the cancellation in dI turns RCPPS's relative error into ~4e-4 m
        # absolute (see the plane test); weights move by that over 3 tau - 5 mm
        far = stable & (np.abs(dI - 0.005) > 1e-3) & (np.abs(dI - 3 * tau) > 1e-3)
        assert np.all(np.abs(wgt - Wt)[far] <= (f + 1) * 1.5e-2)
        assert np.all(np.abs(dist - D)[far] <= 1.5e-3)
    assert (wgt > 0).sum() > 50


# ---------------------------------------------------------------- de-dup
def test_dedup_fixed_point_equations():
    """The processed mask satisfies the defining equations of the reference's
    de-dup (traversal_wrapper_functions.cpp:633-639, 688-690) at every pixel:
    P = valid & !(P_left & key == key_left) & !(P_up & key == key_up)."""
    rng = np.random.default_rng(3)
    h, w = 60, 90
    base = rng.integers(0, 4, (h, w)) * 8
    vx = np.repeat(np.repeat(rng.integers(0, 3, (h // 6 + 1, w // 9 + 1)) * 8, 6, 0), 9, 1)[:h, :w] + (base & 7)
    vy = np.zeros((h, w), np.int32) + rng.integers(0, 8, (h, w))
    vz = (rng.random((h, w)) < 0.1).astype(np.int32) * 8
    valid = (rng.random((h, w)) < 0.9).astype(np.uint8)
    P = pyoracle.dedup(vx, vy, vz, valid).astype(bool)
    key = np.stack([vx & -8, vy & -8, vz & -8], -1)
    for y in range(h):
        for x in range(w):
            el = x > 0 and P[y, x - 1] and np.array_equal(key[y, x - 1], key[y, x])
            eu = y > 0 and P[y - 1, x] and np.array_equal(key[y - 1, x], key[y, x])
            assert P[y, x] == (bool(valid[y, x]) and not el and not eu), (y, x)
    assert 0.2 < P.mean() < 0.9


# ---------------------------------------------------------------- octree
def _fuse(o, sc, R, t, frames, w=320, h=240, maxcam=4.0, seed0=1000):
    for i in frames:
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, w, h, seed=seed0 + i)
        o.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, maxcam)


def _check_tree(o):
    c = o.counts()
    n = c["side"]
    tree = o.tree()
    pos, sc, par = o.leaves()
    seen = np.zeros(len(pos), bool)

    def walk(branch, org, half, anc):
        for co in range(8):
            p = [org[k] + half * ((co >> k) & 1) for k in range(3)]
            leaf = tree[branch + co]
            a = anc
            if leaf < 0xFFFFFFFE:
                assert list(pos[leaf]) == p, "leaf position = its octant origin"
                assert sc[leaf] * 8 == half, "leaf scale = octant size / 8"
                pl = par[leaf]
                if pl != 0xFFFFFFFE:
                    assert pl in anc, "leaf parent is a brick on the path above it"
                seen[leaf] = True
                a = anc + [leaf]
            child = tree[branch + co + 8]
            if child != 0xFFFFFFFE:
                assert child % 16 == 0 and half > 8
                walk(child, p, half // 2, a)

    walk(0, [0, 0, 0], n // 2, [])
    assert seen.all(), "every brick is reachable from the root"
    assert np.all(pos % np.maximum(sc * 8, 1)[:, None] == 0)


def test_tree_invariants_c1():
    sc, R, t, cfg = synth.config_c1(n_frames=100)
    o = Oracle(cfg["scale"], cfg["threshold"], True, 1 << 17, 1 << 21)
    _fuse(o, sc, R, t, range(0, 100, 10))
    _check_tree(o)
    assert o.counts()["n_leaves_used"] > 100


def grow_sequence():
    """First frame close to a wall (a small initial volume), then views across
    the room, which leave the volume and make the tree grow upward."""
    sc = synth.Scene.room_scene(21, (8.0, 8.0, 3.0), 4, 4, 0.1, 0.4)
    n = 8
    R = np.zeros((n, 9)); t = np.zeros((n, 3))
    for i in range(n):
        f = np.array([-1.0, 0.0, 0.0]) if i == 0 else np.array([1.0, 0.3 * (i - 4) / 4, -0.2])
        f /= np.linalg.norm(f)
        x = np.cross(f, [0, 0, 1.0]); x /= np.linalg.norm(x)
        R[i] = np.stack([x, np.cross(f, x), f], 1).ravel()
        t[i] = [0.5 + 0.2 * i, 4.0, 1.5]
    return sc, R, t


def test_grow_keeps_invariants():
    """grow() (geometryfusion_mipmap_cpu.cpp:3071-3319) keeps the tree valid."""
    sc, R, t = grow_sequence()
    o = Oracle(0.02, 0.06, True, 1 << 17, 1 << 21)
    _fuse(o, sc, R, t, [0])
    n0 = o.counts()["side"]
    _fuse(o, sc, R, t, range(1, len(R)), maxcam=6.0)
    c = o.counts()
    assert c["grows"] >= 1 and c["side"] > n0
    _check_tree(o)


# ---------------------------------------------------------------- meshing
def test_mc_table_invariants():
    from tests.mc_reference import EDGE_CORNERS, load_tri_table
    tri = load_tri_table(os.path.join(ROOT, "fastfusion_amd", "csrc", "mc_tables.h"))
    for ci in range(256):
        cut = {e for e, (a, b) in enumerate(EDGE_CORNERS) if ((ci >> a) & 1) != ((ci >> b) & 1)}
        used = set(tri[ci])
        assert used == cut, ci            # exactly the cut edges are used
        assert len(tri[ci]) % 3 == 0
    assert len(tri[0]) == 0 and len(tri[255]) == 0


def test_mc_table_closed_surface():
    """Marching cubes with the table on analytic spheres gives a closed,
    consistently oriented surface: every edge is shared by exactly two
    triangles, traversed once in each direction."""
    from collections import Counter

    from tests.mc_reference import CORNER, EDGE_CORNERS, load_tri_table
    tri = load_tri_table(os.path.join(ROOT, "fastfusion_amd", "csrc", "mc_tables.h"))
    for c0, r in [((6.3, 6.1, 5.8), 3.7), ((7.1, 6.4, 6.9), 4.6), ((6.5, 6.5, 6.5), 2.2)]:
        N = 14
        g = np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij"), -1).astype(float)
        d = np.linalg.norm(g - np.array(c0), axis=-1) - r
        directed = Counter()
        for x in range(N - 1):
            for y in range(N - 1):
                for z in range(N - 1):
                    cs = [(x + a, y + b, z + c) for a, b, c in CORNER]
                    ci = sum(1 << k for k, q in enumerate(cs) if d[q] < 0)
                    row = tri[ci]
                    for t in range(0, len(row), 3):
                        ids = []
                        for e in row[t:t + 3]:
                            a, b = EDGE_CORNERS[e]
                            pa, pb = sorted([cs[a], cs[b]])
                            ids.append((pa, pb))
                        for k in range(3):
                            directed[(ids[k], ids[(k + 1) % 3])] += 1
        for (a, b), cnt in directed.items():
            assert cnt == 1 and directed[(b, a)] == 1, (c0, a, b)


def test_mesh_equals_global_marching_cubes_uniform_scale():
    """All bricks at scale 1 (every depth below 1.125 m): the union of all
    cell meshes equals marching cubes over the global sample grid."""
    from tests.mc_reference import global_mc, load_tri_table, mesh_triangles
    sc = synth.Scene.room_scene(5, (4.0, 4.0, 3.0), 3, 3, 0.1, 0.4)
    f = np.array([-1.0, 0.15, -0.1]); f /= np.linalg.norm(f)
    x = np.cross(f, [0, 0, 1.0]); x /= np.linalg.norm(x)
    Rm = np.stack([x, np.cross(f, x), f], 1).ravel()
    o = Oracle(0.01, 0.03, True, 1 << 16, 1 << 20)
    for i in range(3):
        tt = np.array([0.8, 2.0 + 0.01 * i, 1.5])
        d, bgr = sc.render(Rm.reshape(3, 3), tt, K, 320, 240, seed=i)
        o.add_map(d, Rm, tt, K, bgr, 1 / 5000.0, 1.1)
    pos, s_, _ = o.leaves()
    assert np.all(s_ == 1)
    desc, nv, nf, xyz, rgb, faces = o.mesh_all()
    dd, ww, _ = o.bricks(np.arange(len(pos)))
    tri = load_tri_table(os.path.join(ROOT, "fastfusion_amd", "csrc", "mc_tables.h"))
    A = global_mc(pos, dd, ww, o.counts()["offset"], 0.01, tri)
    B = mesh_triangles(xyz, faces, nv, nf)
    assert len(A) > 5000
    assert A == B


def test_mesh_surface_near_analytic_sdf():
    """Mesh vertices of a fused room lie on the scene surface: the analytic
    SDF at every vertex is within a few voxels (noise + truncation)."""
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    o = Oracle(cfg["scale"], cfg["threshold"], True, 1 << 17, 1 << 21)
    _fuse(o, sc, R, t, range(0, 40, 4))
    desc, nv, nf, xyz, rgb, faces = o.mesh_all()
    assert len(faces) > 2000
    sdf = np.abs(sc.sdf(xyz))
    assert np.median(sdf) < 0.5 * cfg["scale"] * 4
    assert np.percentile(sdf, 95) < 4 * cfg["scale"] * 4


def _triangles(xyz, rgb, faces):
    """Triangles as sorted rows of their corners' (x, y, z bits, r, g, b)."""
    v = np.concatenate([xyz.view(np.uint32).astype(np.int64), rgb.astype(np.int64)], axis=1)
    t = v[faces.astype(np.int64)].reshape(len(faces), 18)
    return t[np.lexsort(t.T[::-1])] if len(t) else t


def _scratch_triangles(o):
    desc, nv, nf, xyz, rgb, faces = o.mesh_all()
    base = np.repeat(np.concatenate([[0], np.cumsum(nv)[:-1]]).astype(np.int64), nf)
    return _triangles(xyz, rgb, faces.astype(np.int64) + base[:, None])


def _incremental_walk(cap):
    """Fuses 12 frames of the 640x480 walk (at 320x240, bricks of scales 1,
    2 and 4) with an incremental update after each; per update, whether the
    incremental mesh equals the from-scratch mesh, and the cells re-meshed."""
    sc, R, t, cfg = synth.config_c2(n_frames=60)
    o = Oracle(cfg["scale"], cfg["threshold"], True, 1 << 17, 1 << 21)
    o.lib.ffo_mesh_set_depth_cap(cap)
    same, remeshed = [], []
    try:
        for i in range(0, 24, 2):
            _fuse(o, sc, R, t, [i])
            remeshed.append(o.update_meshes())
            same.append(np.array_equal(_triangles(*o.mesh_incremental()), _scratch_triangles(o)))
    finally:
        o.lib.ffo_mesh_set_depth_cap(16)
    scales = np.unique(o.leaves()[1])
    return same, remeshed, scales


def test_incremental_mesh_equals_scratch():
    """The incremental update re-meshes a kept cell only when a brick its
    samples read changed (its lastLeafs, and their parent chains up to the
    level its samples reached); after every update of a mixed-scale walk its
    mesh equals the from-scratch mesh of every cell, triangle for triangle
    and bit for bit.  With the check cut to the lastLeafs alone (a planted
    mistake: the parent-chain fallback ignored) some update's mesh differs,
    so the walk exercises the fallback."""
    same, remeshed, scales = _incremental_walk(16)
    assert len(scales) >= 3, scales
    assert all(same), same
    assert min(remeshed[1:]) > 0
    same1, _, _ = _incremental_walk(1)
    assert not all(same1)
