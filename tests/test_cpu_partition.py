"""The incremental cell partition, pinned on the CPU.

The mesh update keeps the cell table between updates and, when the tree
gained only a few bricks and branch blocks since the table was built,
recomputes the cells of the terminal nodes whose option box [p - 1, p + s]
(the voxels the ownership and lastLeaf lookups of a node of origin p and
size s read) meets one of the new bricks' or branches' boxes, and keeps
every other cell, found by its owner (the key's origin plus the recorded
owner offset).  This file restates the partition (terminal nodes and the 27
candidate cells of each, the rules of meshcelltraversalSplitReferenced.hpp
as the oracle restates them), checks the restatement against the oracle's
own cells on every frame, and then checks that kept + recomputed cells equal
the from-scratch partition after every frame of a mixed-scale sequence.  A
planted mistake (the option box one voxel too narrow) must fail.
"""
import numpy as np

from fastfusion_amd import synth
from oracle.pyoracle import Oracle

BR = 4294967294  # empty tree slot
BL = 8


def locate(tree, nl, ns, n, q):
    """Terminal node containing voxel q: size, deepest brick on its path."""
    if min(q) < 0 or max(q) >= n:
        return 0, BR
    half, branch, o, leaf = n >> 1, 0, [0, 0, 0], BR
    while True:
        co = 0
        for k in range(3):
            if q[k] >= o[k] + half:
                co |= 1 << k
                o[k] += half
        lf = int(tree[branch + co])
        if lf < nl:
            leaf = lf
        child = int(tree[branch + co + 8])
        if child >= ns or half <= BL:
            return half, leaf
        branch, half = child, half >> 1


def branch_geometry(tree, ns, n):
    """Origin and octant size of every branch block, by walking the tree."""
    geo = {0: (0, 0, 0, n >> 1)}
    stack = [(0, (0, 0, 0), n >> 1)]
    while stack:
        b, o, half = stack.pop()
        for co in range(8):
            child = int(tree[b + co + 8])
            if child < ns and half > BL:
                p = tuple(o[k] + ((co >> k) & 1) * half for k in range(3))
                geo[child // 16] = (p[0], p[1], p[2], half >> 1)
                stack.append((child, p, half >> 1))
    return geo


def terminals(snap, geo):
    tree, nl, ns, n = snap["tree"], snap["nL"], snap["nS"], snap["n"]
    out = []
    for b, (x0, y0, z0, half) in geo.items():
        for co in range(8):
            if int(tree[b * 16 + 8 + co]) >= ns and half:
                p = (x0 + half * (co & 1), y0 + half * ((co >> 1) & 1), z0 + half * (co >> 2))
                sz, lt = locate(tree, nl, ns, n, p)
                if sz == half and lt != BR:
                    out.append((p, half, lt))
    return out


def node_cells(snap, p, s, lt):
    """Cells owned by terminal node (p, s): ((type, origin, size), lastLeafs, owner offset bits)."""
    tree, nl, ns, n = snap["tree"], snap["nL"], snap["nS"], snap["n"]

    def at(k, c):  # option coordinate c of axis k: p - 1, p, p + s - 1, p + s
        return (p[k] - 1, p[k], p[k] + s - 1, p[k] + s)[c]

    def look(code):
        if all(c in (1, 2) for c in code):
            return s, lt
        return locate(tree, nl, ns, n, tuple(at(k, code[k]) for k in range(3)))

    cells = [((0, p, s), (lt,), 0)]
    for a in range(3):
        for side in (0, 1):
            code = [1, 1, 1]
            code[a] = 3 if side else 0
            nsz, lf = look(code)
            if nsz and (nsz >= s if side else nsz > s):
                o = list(p)
                if not side:
                    o[a] -= s
                ll = (lt, lf) if side else (lf, lt)
                cells.append(((1 + a, tuple(o), s), ll, 0 if side else 1 << a))
    for e in range(3):
        u, v = (1 if e == 0 else 0), (1 if e == 2 else 2)
        for cu in (0, 1):
            for cv in (0, 1):
                mine, ok, ll = (1 - cu) + 2 * (1 - cv), True, []
                for sl in range(4):
                    code = [0, 0, 0]
                    code[e], code[u], code[v] = 1, 2 * cu + (sl & 1), 2 * cv + (sl >> 1)
                    rs, rl = look(code)
                    ll.append(rl)
                    if sl != mine and (not rs or rs < s or (sl < mine and rs == s)):
                        ok = False
                if ok:
                    o = [0, 0, 0]
                    o[e], o[u], o[v] = p[e], p[u] + cu * s - s, p[v] + cv * s - s
                    cells.append(((4 + e, tuple(o), s), tuple(ll), (0 if cu else 1 << u) | (0 if cv else 1 << v)))
    for cb in range(8):
        c3 = (cb & 1, (cb >> 1) & 1, cb >> 2)
        mine, ok, ll = (1 - c3[0]) + 2 * (1 - c3[1]) + 4 * (1 - c3[2]), True, []
        for sl in range(8):
            rs, rl = look((2 * c3[0] + (sl & 1), 2 * c3[1] + ((sl >> 1) & 1), 2 * c3[2] + (sl >> 2)))
            ll.append(rl)
            if sl != mine and (not rs or rs < s or (sl < mine and rs == s)):
                ok = False
        if ok:
            o = tuple(p[k] + c3[k] * s - s for k in range(3))
            cells.append(((7, o, s), tuple(ll), sum((0 if c3[k] else 1) << k for k in range(3))))
    # pushMeshCellQueue meshes a cell only when every side's brick exists
    return [c for c in cells if all(x != BR for x in c[1])]


def partition(snap):
    geo = branch_geometry(snap["tree"], snap["nS"], snap["n"])
    cells = []
    for p, s, lt in terminals(snap, geo):
        cells += node_cells(snap, p, s, lt)
    return cells, geo


def touched(p, s, boxes, margin=1):
    lo = [p[k] - margin for k in range(3)]
    hi = [p[k] + s - 1 + margin for k in range(3)]
    for x, y, z, w in boxes:
        b = (x, y, z)
        if all(lo[k] < b[k] + w and b[k] <= hi[k] for k in range(3)):
            return True
    return False


def incremental(old, new, geo_new, boxes, margin=1):
    """Old cells whose owner is untouched, plus the cells of touched new terminal nodes."""
    kept = []
    for key, ll, own in old:
        _, o, s = key
        owner = tuple(o[k] + (s if (own >> k) & 1 else 0) for k in range(3))
        if not touched(owner, s, boxes, margin):
            kept.append((key, ll, own))
    fresh = []
    for p, s, lt in terminals(new, geo_new):
        if touched(p, s, boxes, margin):
            fresh += node_cells(new, p, s, lt)
    return kept + fresh


def _walk_snapshots(n_frames=12):
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    K = (262.5, 262.5, 159.5, 119.5)
    ora = Oracle(cfg["scale"], cfg["threshold"], True, max_leaves=1 << 17, max_branches=1 << 22)
    snaps, descs = [], []
    for i in range(n_frames):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + i)
        ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        c = ora.counts()
        pos, scl, _ = ora.leaves()
        snaps.append(dict(tree=ora.tree().copy(), nL=c["n_leaves_used"], nS=c["n_branches_used"], n=c["side"],
                          pos=pos.copy(), sc=scl.copy(), grows=c["grows"]))
        descs.append(ora.mesh_all()[0])
    return snaps, descs


def _boxes(a, b, geo_b):
    boxes = [(int(b["pos"][l][0]), int(b["pos"][l][1]), int(b["pos"][l][2]), int(b["sc"][l]) * BL)
             for l in range(a["nL"], b["nL"])]
    boxes += [(geo_b[k][0], geo_b[k][1], geo_b[k][2], 2 * geo_b[k][3]) for k in range(a["nS"] // 16, b["nS"] // 16)]
    return boxes


def _as_set(cells):
    return sorted((key, ll) for key, ll, _ in cells)


def test_incremental_partition_equals_full():
    snaps, descs = _walk_snapshots()
    scales = set()
    checked = planted_fail = 0
    for i, snap in enumerate(snaps):
        full, geo = partition(snap)
        # the restatement equals the oracle's cells (type, origin, size, lastLeafs)
        mine = sorted((k[0], *k[1], k[2], *(list(ll) + [BR] * (8 - len(ll)))) for k, ll, _ in full)
        ref = sorted(tuple(int(v) & 0xffffffff for v in r) for r in descs[i])
        assert mine == [tuple(r) for r in ref], f"frame {i}: restated partition differs from the oracle's"
        scales.update(int(x) for x in snap["sc"])
        if i == 0 or snaps[i - 1]["grows"] != snap["grows"]:
            continue
        old, _ = partition(snaps[i - 1])
        boxes = _boxes(snaps[i - 1], snap, geo)
        assert _as_set(incremental(old, snap, geo, boxes)) == _as_set(full), f"frame {i}: incremental != full"
        # a planted mistake: the option box one voxel too narrow
        planted_fail += _as_set(incremental(old, snap, geo, boxes, margin=0)) != _as_set(full)
        checked += 1
    assert checked >= 8 and len(scales) >= 2  # mixed scales, several incremental steps
    assert planted_fail >= 1
