"""Pins the oracle's math layer (math.glsl RNG, trig definition, disk mapping, half conversion).

The reference ships no golden vectors; these checks use closed forms,
numpy's IEEE implementations and statistical properties.
"""
import ctypes as C

import numpy as np

import orc

L = orc.lib()


def np_hash2(s):
    """Wang integer hash (math.glsl:227-234) on numpy uint32 arrays."""
    s = s.astype(np.uint32)
    s = (s ^ np.uint32(61)) ^ (s >> np.uint32(16))
    s = s * np.uint32(9)
    s = s ^ (s >> np.uint32(4))
    s = s * np.uint32(0x27D4EB2D)
    s = s ^ (s >> np.uint32(15))
    return s


def test_hash2_known_answers():
    # Wang hash known answers, computed by hand from its definition with Python integers
    def wang(x):
        x = ((x ^ 61) ^ (x >> 16)) & 0xFFFFFFFF
        x = (x * 9) & 0xFFFFFFFF
        x = x ^ (x >> 4)
        x = (x * 0x27D4EB2D) & 0xFFFFFFFF
        return x ^ (x >> 15)

    for v in [0, 1, 61, 0xFFFFFFFF, 0x80000000, 123456789]:
        assert L.orc_hash2(v) == wang(v)
    g = np.random.default_rng(0).integers(0, 2**32, 50000, dtype=np.uint64).astype(np.uint32)
    ref = np_hash2(g)
    got = np.array([L.orc_hash2(int(v)) for v in g[:2000]], np.uint32)
    assert np.array_equal(got, ref[:2000])


def test_make_seed_matches_definition():
    # makeSeed(seed, uvec2 index) = hash2((seed + x) ^ (y - 1)) + hash2(y * (x - 2))   (math.glsl:236-246)
    g = np.random.default_rng(1)
    for _ in range(500):
        seed, x, y = (int(v) for v in g.integers(0, 2**32, 3))
        x %= 4096
        y %= 4096
        a = ((seed + x) & 0xFFFFFFFF) ^ ((y - 1) & 0xFFFFFFFF)
        b = (y * ((x - 2) & 0xFFFFFFFF)) & 0xFFFFFFFF
        want = (int(np_hash2(np.array([a]))[0]) + int(np_hash2(np.array([b]))[0])) & 0xFFFFFFFF
        assert L.orc_make_seed(seed, x, y) == want


def test_sample1f_range_and_rounding():
    out = np.zeros(200000, np.float32)
    L.orc_sample1f(12345, out.size, out.ctypes.data)
    assert out.min() >= 0.0 and out.max() <= 1.0
    # uniformity: 20-bin chi-square
    h, _ = np.histogram(out, bins=20, range=(0, 1))
    e = out.size / 20
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 60.0
    # float(u) / 4294967295.0 in FP32: the literal rounds to 2^32 and float(u) rounds up to 2^32
    # for u >= 2^32 - 128, so the result can be exactly 1.0
    u = np.uint32(0xFFFFFFC0)
    assert np.float32(u) / np.float32(4294967295.0) == np.float32(1.0)


def test_sincos_against_libm():
    x = np.linspace(-np.pi, 2 * np.pi, 200001).astype(np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    L.orc_sincos(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
    xs = x.astype(np.float64)
    for got, ref in ((s, np.sin(xs)), (c, np.cos(xs))):
        ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
        err = np.abs(got.astype(np.float64) - ref)
        # within 2 ulp of the correctly rounded value, or 2^-24 absolute near zeros
        assert np.all((err <= 2 * ulp) | (err <= 2 ** -24))


def test_concentric_disk_map():
    uv = np.array([[1.0, 0.5], [0.5, 1.0], [0.0, 0.5], [0.5, 0.0], [0.0, 0.0], [0.5, 0.5]], np.float32)
    out = np.zeros_like(uv)
    L.orc_concentric(uv.ctypes.data, len(uv), out.ctypes.data)
    want = np.array([[1, 0], [0, 1], [-1, 0], [0, -1], [0, 0]], np.float32)
    assert np.allclose(out[:5], want, atol=2e-7)
    # math.glsl:24 tests for zero before remapping to [-1, 1]: the centre (0.5, 0.5)
    # maps to v = 0 and then divides 0 by 0 (a property of the reference, kept)
    assert np.isnan(out[5]).all()
    # Shirley-Chiu mapping preserves area: uniform samples stay uniform on the disk
    g = np.random.default_rng(3)
    uv = g.random((400000, 2)).astype(np.float32)
    out = np.zeros_like(uv)
    L.orc_concentric(uv.ctypes.data, len(uv), out.ctypes.data)
    r2 = (out.astype(np.float64) ** 2).sum(1)
    assert r2.max() <= 1.0 + 1e-6
    h, _ = np.histogram(r2, bins=10, range=(0, 1))  # r^2 uniform for a uniform disk
    e = len(uv) / 10
    assert ((h - e) ** 2 / e).sum() < 40.0
    ang = np.arctan2(out[:, 1], out[:, 0])
    h, _ = np.histogram(ang, bins=16, range=(-np.pi, np.pi))
    e = len(uv) / 16
    assert ((h - e) ** 2 / e).sum() < 60.0


def test_half_conversion_matches_ieee():
    g = np.random.default_rng(4)
    vals = np.concatenate([g.normal(scale=s, size=4000) for s in (1e-6, 1e-4, 1e-2, 1, 100, 7e4)]).astype(np.float32)
    vals = np.concatenate([vals, np.array([0.0, -0.0, 65504.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, np.inf, -np.inf],
                                          np.float32)])
    want = vals.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_f32_to_f16(C.c_float(float(v))) for v in vals], np.uint16)
    assert np.array_equal(got, want)
