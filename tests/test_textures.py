"""Albedo textures (uTextures, layouts.glsl:178; sampled in ray_layouts.glsl:80-86 and
GBuffer.frag:33-37; loaded by Resource::addImage through stb_image as 4-channel 8-bit images and
bound as R8G8B8A8_SRGB with a repeat sampler, src/Scene.cpp:428-445, zvk/core/Memory.cpp:75-92).

* rt_read_png decodes what Python's zlib-backed writer below encodes (compressed deflate streams,
  every row filter, grey / grey+alpha / RGB / RGBA / palette with tRNS), and rejects damaged files.
* The oracle's sampler against what the Vulkan texel-filtering rules fix: the sRGB transfer
  function (IEC 61966-2-1) at the byte values, nearest = the texel containing the coordinate,
  bilinear at a texel centre = that texel exactly, bilinear of a linear ramp = the ramp (within the
  8-bit weight quantisation), the repeat wrap (u and u + 1 sample alike), and the 8-bit weight
  rounding (a weight of 1/512 rounds to even).
* Scene files with map_Kd and <baseColor image=...> load their textures, and the oracle's
  G-buffer shows the texture on a quad facing the camera.
"""
import ctypes as C
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import orc
import rtpkg

rt = rtpkg.rt
L = orc.lib()


def png_bytes(pixels, ctype, palette=None, trns=None, filters=None, level=6):
    """PNG file bytes for an 8-bit image (pixels: H x W x channels uint8; palette index image for
    ctype 3), each row with the given filter type (0-4, cycled)."""
    H, W = pixels.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    px = pixels.reshape(H, W * ch).astype(np.int32)
    filters = filters or [0]
    raw = bytearray()
    for y in range(H):
        f = filters[y % len(filters)]
        row = px[y]
        up = px[y - 1] if y else np.zeros_like(row)
        left = np.concatenate([np.zeros(ch, np.int32), row[:-ch]])
        ul = np.concatenate([np.zeros(ch, np.int32), up[:-ch]])
        if f == 0:
            out = row
        elif f == 1:
            out = row - left
        elif f == 2:
            out = row - up
        elif f == 3:
            out = row - (left + up) // 2
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
            out = row - pred
        raw.append(f)
        raw.extend((out & 0xFF).astype(np.uint8).tobytes())

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, ctype, 0, 0, 0))
    if palette is not None:
        data += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        data += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    data += chunk(b"IDAT", zlib.compress(bytes(raw), level)) + chunk(b"IEND", b"")
    return data


def read_png(path):
    w, h = C.c_uint32(), C.c_uint32()
    rt.check(rt.host().rt_read_png(str(path).encode(), C.byref(w), C.byref(h), None), "rt_read_png size")
    out = np.zeros((h.value, w.value, 4), np.uint8)
    rt.check(rt.host().rt_read_png(str(path).encode(), C.byref(w), C.byref(h), out.ctypes.data), "rt_read_png")
    return out


@pytest.mark.parametrize("ctype", [0, 2, 3, 4, 6])
@pytest.mark.parametrize("level", [0, 1, 9])
def test_png_decoder_round_trip(tmp_path, ctype, level):
    g = np.random.default_rng(ctype * 10 + level)
    W, H = 37, 23
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    # smooth content plus noise so that every filter and both Huffman block types occur
    yy, xx = np.mgrid[0:H, 0:W]
    base = ((xx * 7 + yy * 3)[..., None] + np.arange(ch) * 40) % 256
    pix = ((base + g.integers(0, 3, (H, W, ch))) % 256).astype(np.uint8)
    palette = trns = None
    if ctype == 3:
        palette = g.integers(0, 256, (16, 3))
        trns = g.integers(0, 256, 10)
        pix = g.integers(0, 16, (H, W, 1)).astype(np.uint8)
    path = tmp_path / "t.png"
    path.write_bytes(png_bytes(pix, ctype, palette, trns, filters=[0, 1, 2, 3, 4], level=level))
    got = read_png(path)
    if ctype == 0:
        want = np.concatenate([np.repeat(pix, 3, 2), np.full((H, W, 1), 255, np.uint8)], 2)
    elif ctype == 4:
        want = np.concatenate([np.repeat(pix[..., :1], 3, 2), pix[..., 1:]], 2)
    elif ctype == 2:
        want = np.concatenate([pix, np.full((H, W, 1), 255, np.uint8)], 2)
    elif ctype == 6:
        want = pix
    else:
        idx = pix[..., 0]
        alpha = np.full(16, 255)
        alpha[:10] = trns
        want = np.concatenate([np.asarray(palette)[idx], alpha[idx][..., None]], 2).astype(np.uint8)
    assert np.array_equal(got, want)


def test_png_decoder_rejects_damage(tmp_path):
    pix = np.zeros((4, 4, 4), np.uint8)
    good = bytearray(png_bytes(pix, 6))
    bad_crc = bytearray(good)
    bad_crc[30] ^= 1       # inside IHDR: the chunk CRC no longer matches
    interlaced = png_bytes(pix, 6)
    ihdr = struct.pack(">IIBBBBB", 4, 4, 8, 6, 0, 0, 1)
    interlaced = (interlaced[:8] + struct.pack(">I", 13) + b"IHDR" + ihdr
                  + struct.pack(">I", zlib.crc32(b"IHDR" + ihdr) & 0xFFFFFFFF) + interlaced[33:])
    for name, data in (("crc", bytes(bad_crc)), ("interlaced", interlaced), ("trunc", bytes(good[:40])),
                       ("notpng", b"GIF89a" + bytes(40))):
        p = tmp_path / f"{name}.png"
        p.write_bytes(data)
        w, h = C.c_uint32(), C.c_uint32()
        assert rt.host().rt_read_png(str(p).encode(), C.byref(w), C.byref(h), None) != 0, name
        if name == "interlaced":
            assert b"interlaced" in rt.core().rt_last_error()


Tex = rt.Texture


def sample(texels, filt, uv):
    H, W = texels.shape[:2]
    t = Tex(W, H, filt, 0, texels.ctypes.data)
    uv = np.ascontiguousarray(uv, np.float32)
    out = np.zeros((len(uv), 3), np.float32)
    L.orc_sample_texture(C.byref(t), uv.ctypes.data, len(uv), out.ctypes.data)
    return out


def srgb_table():
    t = np.zeros(256, np.float32)
    L.orc_srgb_table(t.ctypes.data)
    return t


def test_srgb_transfer_function():
    t = srgb_table().astype(np.float64)
    c = np.arange(256) / 255.0
    # IEC 61966-2-1 decoding, float64
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    assert np.array_equal(t, want.astype(np.float32).astype(np.float64))
    assert t[0] == 0.0 and t[255] == 1.0 and (np.diff(t) > 0).all()
    assert abs(t[188] - 0.5) < 3e-3   # sRGB 188 is mid-grey in linear light


def test_sampler_nearest_and_bilinear():
    g = np.random.default_rng(5)
    W, H = 8, 5
    tex = np.ascontiguousarray(g.integers(0, 256, (H, W, 4)).astype(np.uint8))
    lin = srgb_table()[tex[..., :3]]
    # nearest: the texel containing (u, v), with the repeat wrap
    uv = np.stack([g.uniform(-3, 3, 4000), g.uniform(-3, 3, 4000)], 1).astype(np.float32)
    got = sample(tex, 0, uv)
    ix = np.floor(uv[:, 0].astype(np.float64) * W).astype(int) % W
    iy = np.floor(uv[:, 1].astype(np.float64) * H).astype(int) % H
    inside = np.abs(uv[:, 0] * W - np.round(uv[:, 0] * W)) > 1e-4   # away from texel edges (rounding)
    inside &= np.abs(uv[:, 1] * H - np.round(uv[:, 1] * H)) > 1e-4
    assert np.array_equal(got[inside], lin[iy[inside], ix[inside]])
    # bilinear at texel centres returns the texel exactly
    cy, cx = np.mgrid[0:H, 0:W]
    centres = np.stack([(cx.ravel() + 0.5) / W, (cy.ravel() + 0.5) / H], 1)
    got = sample(tex, 1, centres)
    assert np.array_equal(got, lin[cy.ravel(), cx.ravel()])
    # repeat: u + 1 and v - 2 sample alike (coordinates exact in float32)
    uv = (np.stack([g.integers(0, 64, 500), g.integers(0, 64, 500)], 1) / 64.0).astype(np.float32)
    assert np.array_equal(sample(tex, 1, uv), sample(tex, 1, uv + np.float32([1.0, -2.0])))


def test_bilinear_reproduces_a_ramp():
    """A texture whose linear values grow along u by a constant step: bilinear filtering between
    texel centres gives the linear interpolation, up to the 8-bit weight quantisation."""
    lut = srgb_table()
    W, H = 16, 4
    # bytes whose linear values are as evenly spaced as the sRGB curve allows, then the true steps
    b = np.searchsorted(lut, np.linspace(0.05, 0.9, W)).astype(np.uint8)
    tex = np.zeros((H, W, 4), np.uint8)
    tex[..., 0] = tex[..., 1] = tex[..., 2] = b[None, :]
    tex[..., 3] = 255
    vals = lut[b].astype(np.float64)
    u = np.linspace(0.5 / W, (W - 0.5) / W, 301)
    uv = np.stack([u, np.full_like(u, 0.5)], 1).astype(np.float32)
    got = sample(np.ascontiguousarray(tex), 1, uv)[:, 0].astype(np.float64)
    x = uv[:, 0].astype(np.float64) * W - 0.5
    want = np.interp(x, np.arange(W), vals)
    step = np.abs(np.diff(vals)).max()
    assert np.abs(got - want).max() <= step / 512 + 1e-6


def test_bilinear_weight_rounding():
    """Weights carry 8 fractional bits, rounded to nearest even: between a black texel and a white
    one the filtered value is the weight itself; fractions of 1/512, 3/512, 5/512 and 129/512 give
    0, 2/256, 2/256 and 64/256."""
    tex = np.zeros((1, 2, 4), np.uint8)
    tex[0, 1, :3] = 255
    tex[..., 3] = 255
    f = np.array([1 / 512, 3 / 512, 5 / 512, 129 / 512, 0.25], np.float64)
    u = (0.5 + f) / 2   # between the centres of texel 0 and texel 1
    uv = np.stack([u, np.full_like(u, 0.5)], 1).astype(np.float32)
    got = sample(np.ascontiguousarray(tex), 1, uv)[:, 0]
    assert list(got) == [0.0, np.float32(2 / 256), np.float32(2 / 256), np.float32(0.25), np.float32(0.25)]


TEXTURED_XML = """<?xml version="1.0" encoding="utf-8"?>
<scene version="1.0">
  <integrator type="restirpt"><size width="48" height="32"/></integrator>
  <camera type="pinhole">
    <position value="0 -3 1"/><lookAt value="0 0 1"/><fov value="40"/>
    <lensRadius value="0.0"/><focalDistance value="1.0"/>
  </camera>
  <modelInstances>
    <model name="board" path="board.obj">
      <transform translate="0 0 0" scale="1 1 1" rotate="0 0 0"/>
      <material type="lambertian"><baseColor image="checker.png" filter="nearest"/></material>
    </model>
    <model name="floor" path="floor.obj">
      <transform translate="0 0 0" scale="1 1 1" rotate="0 0 0"/>
      <material type="default"/>
    </model>
    <model type="light" name="lamp" path="lamp.obj" flip="true">
      <transform translate="0 0 0" scale="1 1 1" rotate="0 0 0"/>
      <radiance value="20 20 20"/>
    </model>
  </modelInstances>
</scene>
"""
# y-up meshes (the scene turns them z-up): a 2 x 2 board standing at z = 0 (world y = 0) facing the
# camera, UVs over [0, 1]^2; a floor with an MTL texture map; a small ceiling lamp
BOARD_OBJ = """v -1 0 0
v 1 0 0
v 1 2 0
v -1 2 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
f 1/1/1 2/2/1 3/3/1 4/4/1
"""
FLOOR_OBJ = """mtllib floor.mtl
v -3 0 -3
v 3 0 -3
v 3 0 3
v -3 0 3
vt 0 0
vt 4 0
vt 4 4
vt 0 4
vn 0 1 0
usemtl wood
f 1/1/1 4/4/1 3/3/1 2/2/1
"""
FLOOR_MTL = """newmtl wood
Kd 0.5 0.5 0.5
map_Kd ramp.png
"""
LAMP_OBJ = """v -0.3 2.5 -0.3
v 0.3 2.5 -0.3
v 0.3 2.5 0.3
v -0.3 2.5 0.3
f 1 4 3 2
"""


def write_textured_scene(d):
    d = Path(d)
    checker = np.zeros((4, 4, 3), np.uint8)
    yy, xx = np.mgrid[0:4, 0:4]
    checker[(xx + yy) % 2 == 0] = (230, 40, 20)
    checker[(xx + yy) % 2 == 1] = (20, 60, 220)
    (d / "checker.png").write_bytes(png_bytes(checker, 2))
    ramp = np.zeros((8, 8, 4), np.uint8)
    ramp[..., 0] = np.arange(8)[None, :] * 32
    ramp[..., 1] = np.arange(8)[:, None] * 32
    ramp[..., 2] = 128
    ramp[..., 3] = 255
    (d / "ramp.png").write_bytes(png_bytes(ramp, 6, filters=[1, 4]))
    (d / "board.obj").write_text(BOARD_OBJ)
    (d / "floor.obj").write_text(FLOOR_OBJ)
    (d / "floor.mtl").write_text(FLOOR_MTL)
    (d / "lamp.obj").write_text(LAMP_OBJ)
    (d / "scene.xml").write_text(TEXTURED_XML)
    return d / "scene.xml", checker, ramp


def test_textured_scene_loads_and_shows_its_texture(tmp_path):
    path, checker, ramp = write_textured_scene(tmp_path)
    hs = rt.HostScene.load(path)
    a = hs.arrays
    assert a.numTextures == 2
    texs = (Tex * a.numTextures).from_address(a.textures)
    dims = sorted((t.width, t.height, t.filter) for t in texs)
    assert dims == [(4, 4, 0), (8, 8, 1)]        # checker: nearest (the file); map_Kd: linear
    mats = hs.numpy()["materials"]
    assert sorted(int(t) for t in mats["textureIdx"] if t != 0xFFFFFFFF) == [0, 1]
    # the oracle's G-buffer: the board fills the image centre with the checker pattern (sRGB texels
    # decoded to linear, then packed to 8 bits again)
    osc = orc.OracleScene(a)
    W, H = 48, 32
    cam = hs.camera(0, W, H)[0]
    dn, am, _, _ = osc.gbuffer(cam)
    lut = srgb_table()
    cols = set()
    for y in range(10, 22):
        for x in range(16, 32):
            if dn[y, x, 0] == 0:
                continue
            p = int(am[y, x, 0])
            cols.add(tuple((p >> (8 * k)) & 0xFF for k in range(3)))
    want = {tuple(int(np.floor(np.clip(lut[c], 0, 1) * 255 + 0.5)) for c in col)
            for col in ((230, 40, 20), (20, 60, 220))}
    assert want <= cols and len(cols - want) <= 2   # the checker's two colours (the floor may peek in)


def test_texture_arguments_are_checked():
    a = rt.SceneArrays()
    mat = rt.Material()
    mat.textureIdx = 3
    a.materials, a.numMaterials = C.addressof(mat), 1
    tab = (rt.LightTableEntry * 1)()
    a.lightTable, a.numLightTable = C.addressof(tab), 1
    h = C.c_void_p()
    assert rt.core().rt_device_scene_create(C.byref(a), None, C.byref(h)) == 1
    assert b"textureIdx out of range" in rt.core().rt_last_error()
    t = (Tex * 1)(Tex(0, 4, 1, 0, None))
    a.textures, a.numTextures = C.addressof(t), 1
    mat.textureIdx = 0
    assert rt.core().rt_device_scene_create(C.byref(a), None, C.byref(h)) == 1
    assert b"texture 0" in rt.core().rt_last_error()
