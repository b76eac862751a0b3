"""The reference's raw map format, restated for the checks -- test
infrastructure only (tests/ import it; the product never does).

FusionMipMapCPU::copyCPU2HDD / copyHDD2CPU (geometryfusion_mipmap_cpu.cpp:
1135-1184) write, in this order and without padding: u32 nBranchesUsed,
u32 nLeavesUsed, u32 tree[nBranchesUsed], u16 leafPos[nLeavesUsed][3]
(sidetype3), u16 leafScale[nLeavesUsed] (sidetype), f32 distance[nL][512],
f32 weights[nL][512] (weighttype = float, definitions.h:37-39), and with color
u16 color[nL][512][3] (colortype3).  The file name is getDescription()
(:1128-1133): "%.10f_%.10f_%.10f_%.10f_%.7i" of offset x, y, z, scale, n.
The library appends "FFPARENT", a u32 count and the u32 parents; parents()
reads that block.
"""
import numpy as np

PARENT_TAG = b"FFPARENT"


def description(offset, scale, n):
    return "%.10f_%.10f_%.10f_%.10f_%.7i" % (float(np.float32(offset[0])), float(np.float32(offset[1])),
                                               float(np.float32(offset[2])), float(np.float32(scale)), int(n))


def encode(tree, pos, scale, dist, wgt, color=None, parents=None):
    """Bytes of a map: tree (nB,) u32, pos (nL, 3) u16, scale (nL,) u16,
    dist / wgt (nL, 512) f32, color (nL, 512, 3) u16 or None."""
    tree = np.ascontiguousarray(tree, "<u4")
    pos = np.ascontiguousarray(pos, "<u2").reshape(-1, 3)
    nl = len(pos)
    parts = [np.array([len(tree), nl], "<u4").tobytes(), tree.tobytes(), pos.tobytes(),
             np.ascontiguousarray(scale, "<u2").tobytes(),
             np.ascontiguousarray(dist, "<f4").reshape(nl, 512).tobytes(),
             np.ascontiguousarray(wgt, "<f4").reshape(nl, 512).tobytes()]
    if color is not None:
        parts.append(np.ascontiguousarray(color, "<u2").reshape(nl, 512, 3).tobytes())
    if parents is not None:
        parts += [PARENT_TAG, np.array([nl], "<u4").tobytes(), np.ascontiguousarray(parents, "<u4").tobytes()]
    return b"".join(parts)


def decode(data, color):
    """Inverse of encode (the parents block, when present, as 'parents')."""
    nb, nl = np.frombuffer(data, "<u4", 2)
    o = 8
    out = {}

    def take(dtype, count, shape):
        nonlocal o
        a = np.frombuffer(data, dtype, count, o).reshape(shape)
        o += a.nbytes
        return a

    out["tree"] = take("<u4", nb, (nb,))
    out["pos"] = take("<u2", 3 * nl, (nl, 3))
    out["scale"] = take("<u2", nl, (nl,))
    out["dist"] = take("<f4", 512 * nl, (nl, 512))
    out["wgt"] = take("<f4", 512 * nl, (nl, 512))
    if color:
        out["color"] = take("<u2", 3 * 512 * nl, (nl, 512, 3))
    out["ref_bytes"] = o  # the part a reader of the reference format consumes
    if data[o:o + 8] == PARENT_TAG:
        o += 8
        cnt = int(np.frombuffer(data, "<u4", 1, o)[0])
        o += 4
        out["parents"] = take("<u4", cnt, (cnt,))
    out["end"] = o
    return out


def from_oracle(o, with_parents=True):
    """The oracle's map in the format, from its arrays."""
    n = o.lib.ffo_n_leaves(o.h)
    pos, sc, par = o.leaves()
    d, w, c = o.bricks(np.arange(n))
    col = np.ascontiguousarray(c.transpose(0, 2, 1)) if o.color else None
    return encode(o.tree(), pos, sc, d, w, col, par if with_parents else None)
