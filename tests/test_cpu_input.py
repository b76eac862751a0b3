"""The input side without a GPU: the library's PNG reader against an
independent encoder (Pillow, and a row-filter encoder written here for every
PNG filter type), and TUM association parsing against closed forms
(onlinefusionviewer_main.cpp:266-346: quaternion order qx qy qz qw, image
paths prefixed with the file's directory, poses of a further file composed
with the last pose of the previous one)."""
import struct
import zlib

import numpy as np
import pytest

from fastfusion_amd import FFError, tum

PIL = pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("shape", [(1, 1), (7, 13), (240, 320), (300, 517)])
def test_depth_png_16bit(tmp_path, shape):
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    d = rng.integers(0, 65536, shape).astype(np.uint16)
    d[:, ::5] = 0
    p = str(tmp_path / "d.png")
    PIL.fromarray(d).save(p)  # uint16 -> mode I;16, a 16-bit grayscale PNG
    assert tum.png_size(p) == (shape[1], shape[0], 1, 16)
    assert np.array_equal(tum.read_depth_png(p), d)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P"])
def test_color_png_as_bgr(tmp_path, mode):
    rng = np.random.default_rng(7)
    rgb = rng.integers(0, 256, (61, 83, 3)).astype(np.uint8)
    img = PIL.fromarray(rgb, "RGB")
    if mode == "RGBA":
        a = rng.integers(0, 256, (61, 83)).astype(np.uint8)
        img = PIL.fromarray(np.dstack([rgb, a]), "RGBA")
    elif mode in ("L", "LA", "P"):
        img = img.convert(mode)
    p = str(tmp_path / f"c_{mode}.png")
    img.save(p, compress_level=9)
    want = np.asarray(img.convert("RGB"))[:, :, ::-1]  # imread: B, G, R, alpha dropped
    assert np.array_equal(tum.read_color_png(p), want)


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xffffffff)


def _encode(img, bitdepth, ctype, filters):
    """A PNG whose row y uses filter type filters[y % len] (the specification's
    forward filters)."""
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    raw = img.astype(">u2").tobytes() if bitdepth == 16 else img.astype(np.uint8).tobytes()
    stride = w * ch * bitdepth // 8
    bpp = ch * bitdepth // 8
    rows = [np.frombuffer(raw[y * stride:(y + 1) * stride], np.uint8).astype(np.int32) for y in range(h)]
    out = bytearray()
    for y in range(h):
        x = rows[y]
        up = rows[y - 1] if y else np.zeros_like(x)
        a = np.concatenate([np.zeros(bpp, np.int32), x[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), up[:-bpp]])
        ft = filters[y % len(filters)]
        if ft == 0:
            f = x
        elif ft == 1:
            f = x - a
        elif ft == 2:
            f = x - up
        elif ft == 3:
            f = x - (a + up) // 2
        else:
            p = a + up - c
            pa, pb, pc = np.abs(p - a), np.abs(p - up), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, up, c))
            f = x - pred
        out.append(ft)
        out += (f & 0xff).astype(np.uint8).tobytes()
    z = zlib.compress(bytes(out), 6)
    ihdr = struct.pack(">IIBBBBB", w, h, bitdepth, ctype, 0, 0, 0)
    # the image data split over two IDAT chunks
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", z[:len(z) // 2]) +
            _chunk(b"IDAT", z[len(z) // 2:]) + _chunk(b"IEND", b""))


@pytest.mark.parametrize("filters", [[0], [1], [2], [3], [4], [0, 1, 2, 3, 4]])
def test_every_row_filter(tmp_path, filters):
    rng = np.random.default_rng(11)
    d = rng.integers(0, 65536, (23, 37)).astype(np.uint16)
    d[5:9] = 1234  # flat rows: large residuals only at edges
    p = tmp_path / "d.png"
    p.write_bytes(_encode(d, 16, 0, filters))
    assert np.array_equal(tum.read_depth_png(str(p)), d)
    rgb = rng.integers(0, 256, (19, 29, 3)).astype(np.uint8)
    q = tmp_path / "c.png"
    q.write_bytes(_encode(rgb, 8, 2, filters))
    assert np.array_equal(tum.read_color_png(str(q)), rgb[:, :, ::-1])
    rgb16 = rng.integers(0, 65536, (9, 11, 3)).astype(np.uint16)
    r = tmp_path / "c16.png"
    r.write_bytes(_encode(rgb16, 16, 2, filters))
    assert np.array_equal(tum.read_color_png(str(r)), (rgb16 >> 8).astype(np.uint8)[:, :, ::-1])


def test_png_errors(tmp_path):
    d = np.arange(12, dtype=np.uint16).reshape(3, 4)
    good = _encode(d, 16, 0, [0])
    bad = bytearray(good)
    bad[40] ^= 0xff  # inside a chunk: its CRC no longer matches
    p = tmp_path / "bad.png"
    p.write_bytes(bytes(bad))
    with pytest.raises(FFError, match="CRC"):
        tum.read_depth_png(str(p))
    q = tmp_path / "text.png"
    q.write_bytes(b"not a png")
    with pytest.raises(FFError, match="not a PNG"):
        tum.read_depth_png(str(q))
    rgb = tmp_path / "rgb.png"
    rgb.write_bytes(_encode(np.zeros((2, 2, 3), np.uint8), 8, 2, [0]))
    with pytest.raises(FFError, match="single-channel"):
        tum.read_depth_png(str(rgb))


def _qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _rotate(q, v):
    """q v q* with Hamilton products (q = w, x, y, z)."""
    qc = q * np.array([1, -1, -1, -1])
    return _qmul(_qmul(q, np.concatenate([[0.0], v])), qc)[1:]


def _line(stamp, t, q, depth, rgb=None):
    s = f"{stamp:.6f} {t[0]:.6f} {t[1]:.6f} {t[2]:.6f} {q[0]:.7f} {q[1]:.7f} {q[2]:.7f} {q[3]:.7f} {stamp:.6f} {depth}"
    return s + (f" {stamp:.6f} {rgb}" if rgb else "")


def test_tum_quaternion_order_and_paths(tmp_path):
    rng = np.random.default_rng(2)
    qs = [np.array([0, 0, 0, 1.0]), np.array([0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)])]
    for _ in range(6):
        q = rng.normal(size=4)
        qs.append(q / np.linalg.norm(q))
    ts = rng.normal(size=(len(qs), 3))
    d = tmp_path / "seq"
    d.mkdir()
    lines = [_line(1.0 + i, ts[i], qs[i], f"depth/{i}.png", f"rgb/{i}.png") for i in range(len(qs))]
    (d / "associate.txt").write_text("\n".join(lines) + "\n")
    fr = tum.read_associations([str(d / "associate.txt")])
    assert len(fr) == len(qs)
    for i, (R, t, dp, rp) in enumerate(fr):
        assert dp == str(d) + f"/depth/{i}.png" and rp == str(d) + f"/rgb/{i}.png"
        # the file's numbers are read as floats (float q1..q4, translation1..3)
        qf = np.array([float(np.float32(f"{x:.7f}")) for x in qs[i]])
        assert np.array_equal(t, [float(np.float32(f"{x:.6f}")) for x in ts[i]])
        qh = np.array([qf[3], qf[0], qf[1], qf[2]])  # Quaterniond(q4, q1, q2, q3)
        qh /= np.linalg.norm(qh)  # the file's 7 decimals leave |q| = 1 to ~1e-7
        for v in np.eye(3):
            assert np.allclose(R @ v, _rotate(qh, v), atol=1e-6)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert np.array_equal(fr[0][0], np.eye(3))
    assert np.allclose(fr[1][0], [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)


def test_tum_without_color_blank_lines_and_two_files(tmp_path):
    rng = np.random.default_rng(4)
    q = rng.normal(size=(5, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(size=(5, 3))
    a = tmp_path / "a.txt"
    a.write_text(_line(1, t[0], q[0], "d0.png") + "\n\n" + _line(2, t[1], q[1], "d1.png") + "\n")
    b = tmp_path / "sub"
    b.mkdir()
    (b / "b.txt").write_text("\n".join(_line(3 + i, t[2 + i], q[2 + i], f"d{2 + i}.png") for i in range(3)))
    one = tum.read_associations([str(a)], use_color=False)
    assert [f[2] for f in one] == [str(tmp_path) + "/d0.png", str(tmp_path) + "/d1.png"]
    assert all(f[3] == "" for f in one)
    two = tum.read_associations([str(a), str(b / "b.txt")], use_color=False)
    assert len(two) == 5
    lastR, lastT = one[-1][0], one[-1][1]
    alone = tum.read_associations([str(b / "b.txt")], use_color=False)
    for (R, tt, _, _), (R1, t1, _, _) in zip(two[2:], alone):
        assert np.allclose(R, lastR @ R1, atol=1e-12) and np.allclose(tt, lastR @ t1 + lastT, atol=1e-12)


def test_tum_malformed_line_repeats_previous(tmp_path):
    """A line that does not start with a number leaves the reader's variables
    as they were (operator>> fails from its first extraction), so it repeats
    the previous pose and names, as in the reference."""
    q = np.array([0.1, 0.2, 0.3, 0.9])
    q /= np.linalg.norm(q)
    p = tmp_path / "a.txt"
    p.write_text(_line(1, [1, 2, 3], q, "d0.png", "c0.png") + "\n# a comment line\n")
    fr = tum.read_associations([str(p)])
    assert len(fr) == 2
    assert np.array_equal(fr[0][0], fr[1][0]) and np.array_equal(fr[0][1], fr[1][1])
    assert fr[0][2] == fr[1][2] and fr[0][3] == fr[1][3]
