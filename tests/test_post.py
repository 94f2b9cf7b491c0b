"""Post-processing (post_proc.frag) and the screenshot PNG (Renderer::writeScreenshot): the oracle's
tone curves against their closed forms and fixed points, the texel mapping against its defining
properties, the 8-bit store, and the PNG writer against an independent decoder (zlib + chunk CRCs)."""
import os
import struct
import zlib

import numpy as np
import pytest

import orc
import rtpkg

rt = rtpkg.rt


def filmic64(c):
    f = lambda x: (x * (x * 0.22 + 0.03) + 0.002) / (x * (x * 0.22 + 0.3) + 0.06) - 1.0 / 30.0
    return f(c * 1.6) / f(11.2)


def aces64(c):
    return (c * (c * 2.51 + 0.03)) / (c * (c * 2.43 + 0.59) + 0.14)


def test_tone_curves():
    L = orc.lib()
    xs = np.concatenate([np.linspace(0, 4, 101), [7.0, 20.0, 100.0]])
    for x in xs:
        assert abs(L.orc_tonemap(float(x), 0) - np.float32(x)) == 0
        assert abs(L.orc_tonemap(float(x), 1) - filmic64(x)) < 2e-6 * max(1.0, filmic64(x))
        assert abs(L.orc_tonemap(float(x), 2) - aces64(x)) < 2e-6 * max(1.0, aces64(x))
    # Filmic's white point: c * 1.6 = 11.2 maps to exactly 1; ACES(0) = 0 and ACES(inf) -> 2.51 / 2.43
    assert L.orc_tonemap(7.0, 1) == 1.0
    assert L.orc_tonemap(0.0, 2) == 0.0
    assert abs(L.orc_tonemap(1e6, 2) - 2.51 / 2.43) < 1e-5


def test_texel_mapping_and_store():
    """Pixel (x, y) shows texel floor((W - 1)(x + 0.5) / W): never the last column, monotone, at most one
    pixel behind; gamma 1/2.2 of 0.5 stores round(0.7297 * 255) = 186; values clamp to [0, 1]."""
    H, W = 5, 9
    direct = np.zeros((H, W, 4), np.float32)
    indirect = np.zeros((H, W, 4), np.float32)
    # texel (x, y) encodes its own coordinates in linear values k / 255 (no tone map, no gamma)
    for y in range(H):
        for x in range(W):
            direct[y, x, 0] = x / 255.0
            direct[y, x, 1] = y / 255.0
            indirect[y, x, 2] = 7 / 255.0
    out = orc.post_process(direct, indirect, (0, 0, 0, 0))
    cx, cy = out[..., 0].astype(int), out[..., 1].astype(int)
    for y in range(H):
        row = cx[y]
        assert (np.diff(row) >= 0).all() and row.max() == W - 2 and (np.arange(W) - row <= 1).all()
        assert (row == np.floor((W - 1) * (np.arange(W) + 0.5) / W)).all()
    assert (cy[:, 0] == np.floor((H - 1) * (np.arange(H) + 0.5) / H)).all()
    assert (out[..., 2] == 7).all() and (out[..., 3] == 255).all()
    # switches
    assert (orc.post_process(direct, indirect, (0, 0, 1, 0))[..., :2] == 0).all()
    assert (orc.post_process(direct, indirect, (0, 0, 0, 1))[..., 2] == 0).all()
    grey = np.full((2, 2, 4), 0.5, np.float32)
    z = np.zeros_like(grey)
    assert (orc.post_process(grey, z, (0, 1, 0, 0))[..., :3] == 186).all()
    hot = np.full((2, 2, 4), 50.0, np.float32)
    assert (orc.post_process(hot, z, (0, 0, 0, 0))[..., :3] == 255).all()


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[0][1][:10])
    assert depth == 8 and ctype == 6
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (300, 120)])
def test_png_writer_round_trip(tmp_path, w, h):
    img = np.random.default_rng(w * h).integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    path = str(tmp_path / "shot.png")
    assert rt.host().rt_write_png(path.encode(), img.ctypes.data, w, h) == 0
    assert np.array_equal(read_png(path), img)
    assert rt.host().rt_write_png(path.encode(), None, w, h) != 0
