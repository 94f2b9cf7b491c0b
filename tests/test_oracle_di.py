"""Pins the oracle's ReSTIR DI restatement (TestReSTIR: di_path_gen / di_temporal / di_spatial
over di_reservoir.glsl) against what the mathematics fixes:
  - without reuse, path generation + shading is an unbiased estimator of direct lighting for each
    sample type (light, BSDF, both with MIS), so the mean over seeds matches the closed-form form
    factor of a Lambert floor under a quad light;
  - reusing a pixel's own reservoir through either shift maps the sample onto itself, so its target
    value equals the weight it was resampled with: resampleWeight doubles, the count becomes 2 and
    the reconnection shift recomputes the sample's pdf to the value it was drawn with;
  - the Hybrid shift merges nothing, and the clear flag turns temporal reuse off."""
import ctypes

import numpy as np
import pytest

import orc
import rtpkg
from test_oracle_passes import irradiance_from_quad

rt = rtpkg.rt


def still_pair(cam):
    """uCamera = uPrevCamera = cam (no motion)."""
    c = (rt.Camera * 2)()
    ctypes.memmove(ctypes.addressof(c), ctypes.addressof(cam), ctypes.sizeof(rt.Camera))
    ctypes.memmove(ctypes.addressof(c) + ctypes.sizeof(rt.Camera), ctypes.addressof(cam), ctypes.sizeof(rt.Camera))
    return c


def seeded(cam, f):
    """cam with the seed of frame f (frameIndex 0, so each pass output is that frame's estimate)."""
    c = still_pair(cam)
    h = (f * 0x9E3779B1 + 0x7F4A7C15) & 0xFFFFFFFF
    h ^= h >> 16
    c[0].seed = c[1].seed = (h * 0x85EBCA6B) & 0xFFFFFFFF
    c[0].frameIndex = c[1].frameIndex = 0x80000000
    return c


def di_frame(osc, cams, dn, am, mv, dn_prev, am_prev, s, resv_prev, W, H):
    """path gen -> temporal -> spatial of one frame; returns (this, temp, direct)."""
    n = W * H
    this = orc.reset_di_reservoirs(n)
    temp = orc.reset_di_reservoirs(n)
    out = np.zeros((H, W, 4), np.float32)
    osc.di_path_gen(cams, dn, am, s, this)
    osc.di_temporal(cams, dn, am, mv, dn_prev, am_prev, s, this, resv_prev, temp)
    osc.di_spatial(cams, dn, am, s, temp, this, out)
    return this, temp, out


@pytest.mark.parametrize("sample,frames,tol", [(0, 48, 0.02), (1, 160, 0.03), (2, 48, 0.02)])
def test_di_without_reuse_matches_form_factor(sample, frames, tol):
    hs = rt.HostScene(2, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 24, 24
    n = W * H
    s = rt.di_settings(shift=0, sample=sample, temporal=False, spatial=False)
    acc = np.zeros((H, W, 3))
    base = hs.camera(0, W, H)[0]
    dn, am, mv, _ = osc.gbuffer(base)
    for f in range(frames):
        cams = seeded(base, f)
        _, temp, out = di_frame(osc, cams, dn, am, mv, dn, am, s, orc.reset_di_reservoirs(n), W, H)
        acc += out[..., :3]
        # without reuse the shaded luminance is the reservoir's weight sum (one candidate per sample)
        lum = out[..., :3].astype(np.float64) @ np.array([0.299, 0.587, 0.114])
        m = (temp["instanceIdx"] != 0xFFFFFFFF).reshape(H, W)
        assert np.allclose(lum[m], temp["resampleWeight"].reshape(H, W)[m], rtol=1e-5)
        assert (temp["sampleCount"] == 1).all() or not m.any()
    est = acc / frames
    cam = hs.camera(0, W, H)[0]
    albedo = round(0.5 * 255) / 255.0
    corners = [np.array(c, float) for c in ([-0.5, -0.5, 1.0], [0.5, -0.5, 1.0], [0.5, 0.5, 1.0], [-0.5, 0.5, 1.0])]
    dirs = np.zeros((H, W, 3), np.float32)
    orc.lib().orc_camera_rays(ctypes.byref(cam), dirs.ctypes.data)
    want = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            if dn[y, x, 0] == 0:
                continue
            p = np.array(list(cam.pos)) + dirs[y, x] * (dn[y, x, 0] - 1e-4)
            want[y, x] = albedo / np.pi * irradiance_from_quad(p, corners, 4.0, np.array([0, 0, -1.0]))
    m = dn[..., 0] > 0
    rel = est[..., 0][m].mean() / want[m].mean() - 1.0
    assert abs(rel) < tol, f"sample type {sample}: image-mean direct radiance off by {rel:.4f}"


@pytest.mark.parametrize("shift", [0, 1])
@pytest.mark.parametrize("sample", [0, 1, 2])
def test_self_reuse_doubles_the_weight(shift, sample):
    # BSDF sampling rarely reaches the box room's small light: use the floor under a large quad light
    hs = rt.HostScene(2 if sample == 1 else 0, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 48, 32
    n = W * H
    cams = still_pair(hs.camera(2, W, H)[0])
    assert cams[0].frameIndex & 0x80000000 == 0
    dn, am, _, _ = osc.gbuffer(cams[0])
    mv = np.zeros((H, W, 2), np.uint16)
    s = rt.di_settings(shift=shift, sample=sample, temporal=True, spatial=False)
    this = orc.reset_di_reservoirs(n)
    osc.di_path_gen(cams, dn, am, s, this)
    temp = orc.reset_di_reservoirs(n)
    # previous frame = this frame (same G-buffer, camera and reservoirs)
    osc.di_temporal(cams, dn, am, mv, dn, am, s, this, this, temp)
    fg = dn.reshape(n, 4)[:, 0] > 0
    valid = fg & (this["instanceIdx"] != 0xFFFFFFFF) & (this["resampleWeight"] > 0)
    assert valid.sum() >= 40
    assert (temp["sampleCount"][valid] == 2).all()
    w0, w1 = this["resampleWeight"][valid].astype(np.float64), temp["resampleWeight"][valid].astype(np.float64)
    kept = w1 > 0  # the final visibility ray may lose a sample that grazes an edge
    assert kept.mean() > 0.99
    assert np.allclose(w1[kept], 2.0 * w0[kept], rtol=2e-3), np.abs(w1[kept] / w0[kept] - 2).max()
    # the pixel keeps its own sample (both candidates are that sample)
    if shift == 0:
        assert np.array_equal(temp["triangleIdx"][valid], this["triangleIdx"][valid])
        assert np.array_equal(temp["instanceIdx"][valid], this["instanceIdx"][valid])
        pdf0, pdf1 = this["samplePdf"][valid], temp["samplePdf"][valid]
        assert np.allclose(pdf1, pdf0, rtol=2e-3), np.abs(pdf1 / pdf0 - 1).max()
    # pixels without a sample: reconnection merges a zero-weight candidate; replay's candidate is
    # drawn from the empty sample's zero seed and has weight 0 * pHat / 0 = NaN, so it is dropped
    empty = fg & ~valid
    assert (temp["resampleWeight"][empty] == 0).all()
    if shift == 0:
        assert (temp["sampleCount"][empty] == 2).all()
    else:
        assert np.isin(temp["sampleCount"][empty], (1, 2)).all()


def test_hybrid_shift_and_clear_flag_merge_nothing():
    hs = rt.HostScene(0, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 40, 24
    n = W * H
    cam2 = hs.camera(2, W, H)[0]
    dn, am, _, _ = osc.gbuffer(cam2)
    mv = np.zeros((H, W, 2), np.uint16)
    for shift, clear in ((2, False), (0, True)):
        cams = still_pair(cam2)
        if clear:
            cams[0].frameIndex |= 0x80000000
        s = rt.di_settings(shift=shift, sample=2, temporal=True, spatial=False)
        this = orc.reset_di_reservoirs(n)
        osc.di_path_gen(cams, dn, am, s, this)
        temp = orc.reset_di_reservoirs(n)
        osc.di_temporal(cams, dn, am, mv, dn, am, s, this, this, temp)
        fg = dn.reshape(n, 4)[:, 0] > 0
        assert (temp["sampleCount"][fg] == 1).all()
        same = temp["resampleWeight"][fg] == this["resampleWeight"][fg]
        assert same.mean() > 0.99 and (temp["resampleWeight"][fg][~same] == 0).all()
        # background pixels are not written by any pass
        assert (temp["instanceIdx"][~fg] == 0xFFFFFFFF).all() and (temp["sampleCount"][~fg] == 0).all()


def test_spatial_reuse_is_bounded_by_the_cap():
    """Ten neighbours of one frame plus the temporal reservoir: M never exceeds the cap of 40, and the
    shaded output is finite and non-negative."""
    hs = rt.HostScene(0, 1)
    osc = orc.OracleScene(hs.arrays)
    W, H = 40, 32
    n = W * H
    prev = orc.reset_di_reservoirs(n)
    dn_p = am_p = None
    for f in range(4):
        cams = hs.camera(f, W, H)
        dn, am, mv, _ = osc.gbuffer(cams[0])
        s = rt.di_settings(shift=f % 2, sample=2, temporal=True, spatial=True)
        this, temp, out = di_frame(osc, cams, dn, am, mv, dn if dn_p is None else dn_p,
                                   am if am_p is None else am_p, s, prev, W, H)
        assert this["sampleCount"].max() <= 40 and temp["sampleCount"].max() <= 40
        assert np.isfinite(out).all() and (out[..., :3] >= 0).all()
        prev, dn_p, am_p = this, dn, am
    assert this["sampleCount"].max() > 10
