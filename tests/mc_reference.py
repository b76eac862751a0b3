"""Brute-force marching cubes over the global voxel grid of a uniform-scale
map, used to pin the cell decomposition of the mesh oracle.

With every brick at scale 1, the reference's mesh cells (interior, wall,
edge, corner) tile the dual grid exactly, so the union of all cell meshes
must equal plain marching cubes over all unit cubes whose 8 corner samples
exist, with the zero-weight veto (src/fusion/mesh.hpp:259-291) and the
double-precision VertexInterp (src/fusion/mesh.cpp:1266-1339).  Cubes whose
lower corner is a brick's last sample in all three axes are corner cells,
which the reference meshes non-indexed with Bourke's corner pairs and
without degenerate triangles (mesh_interleaved_meshcell.cpp:2480-2679).
Triangulation: the standard Lorensen-Cline / Bourke table, read here from
its packed form only to reuse the data.
"""
import numpy as np

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE_CORNERS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def load_tri_table(path):
    src = open(path).read()
    body = src[src.index("{") + 1:src.rindex("}")]
    rows = [s.strip().strip('"') for s in body.split(",") if s.strip()]
    assert len(rows) == 256
    return [[int(ch, 16) for ch in r] for r in rows]


def vinterp(p1, p2, v1, v2):
    v1 = float(v1); v2 = float(v2)
    if abs(0.0 - v1) < 1e-40:
        return tuple(p1)
    if abs(0.0 - v2) < 1e-40:
        return tuple(p2)
    if abs(v1 - v2) < 1e-40:
        return tuple(p1)
    mu = (0.0 - v1) / (v2 - v1)
    return tuple(np.float32(float(a) + mu * (float(b) - float(a))) for a, b in zip(p1, p2))


def global_mc(pos, dist, wgt, offset, scale, tri):
    """pos (n, 3) brick origins (scale 1), dist/wgt (n, 512).  Returns the
    multiset (sorted list) of triangles as 9-tuples of float32 coordinates."""
    grid = {}
    for b in range(len(pos)):
        for i in range(512):
            x, y, z = i & 7, (i >> 3) & 7, i >> 6
            grid[(int(pos[b][0]) + x, int(pos[b][1]) + y, int(pos[b][2]) + z)] = (dist[b][i], wgt[b][i])
    off = [np.float32(o) for o in offset]
    sc = np.float32(scale)

    def P(g):
        return tuple(np.float32(off[k] + np.float32(g[k]) * sc) for k in range(3))

    tris = []
    for g in grid:
        cs = [(g[0] + c[0], g[1] + c[1], g[2] + c[2]) for c in CORNER]
        if not all(c in grid for c in cs):
            continue
        vals = [grid[c] for c in cs]
        if any(w == 0.0 for _, w in vals):
            continue
        ci = sum(1 << k for k, (d, _) in enumerate(vals) if d < 0.0)
        row = tri[ci]
        if not row:
            continue
        corner_cell = all((g[k] % 8) == 7 for k in range(3))
        if corner_cell:
            # getColoredTrianglesForCubePrecomputed: corner k at base + bit * (scale * stretch)
            base = P(g)
            pts = [tuple(np.float32(base[k] + np.float32(c[k]) * sc) for k in range(3)) for c in CORNER]
        else:
            pts = [P(c) for c in cs]
        for t in range(0, len(row), 3):
            vs = []
            for e in row[t:t + 3]:
                a, b = EDGE_CORNERS[e]
                if not corner_cell:
                    # indexed cells interpolate every edge from its lower sample
                    if sum(cs[b]) < sum(cs[a]):
                        a, b = b, a
                vs.append(vinterp(pts[a], pts[b], vals[a][0], vals[b][0]))
            if corner_cell and (vs[0] == vs[1] or vs[1] == vs[2] or vs[0] == vs[2]):
                continue
            tris.append(tuple(float(c) for v in vs for c in v))
    return sorted(tris)


def mesh_triangles(xyz, faces, nv, nf):
    """Triangles of a per-cell mesh (faces local to each cell) as sorted 9-tuples."""
    out = []
    v0 = f0 = 0
    for c in range(len(nv)):
        for f in faces[f0:f0 + nf[c]]:
            out.append(tuple(float(xyz[v0 + int(i)][k]) for i in f for k in range(3)))
        v0 += int(nv[c])
        f0 += int(nf[c])
    return sorted(out)
