"""Pins the oracle's per-pixel passes against closed forms and reduction identities."""
import ctypes

import numpy as np
import pytest

import orc
import rtpkg

rt = rtpkg.rt


def irradiance_from_quad(p, corners, radiance, n_light, samples=400):
    """E(p) = int L cos_p cos_l / r^2 dA over a planar quad light (midpoint rule, float64)."""
    a, b, c, d = corners
    u = (np.arange(samples) + 0.5) / samples
    U, V = np.meshgrid(u, u)
    q = a[None, None] + U[..., None] * (b - a)[None, None] + V[..., None] * (d - a)[None, None]
    area = np.linalg.norm(np.cross(b - a, d - a))
    r = q - p
    dist2 = (r ** 2).sum(-1)
    w = r / np.sqrt(dist2)[..., None]
    cos_p = np.maximum(w[..., 2], 0.0)                   # receiver normal +z
    cos_l = np.maximum(-(w * n_light).sum(-1), 0.0)      # emitter faces -z
    return radiance * (cos_p * cos_l / dist2).mean() * area


def test_direct_illumination_matches_form_factor():
    """di_naive's NEE + BSDF-sample RIS with MIS (di_naive.glsl:79-167) is unbiased:
    the mean over seeds converges to albedo/pi * E for a Lambert floor under a quad light."""
    hs = rt.HostScene(2, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 24, 24
    frames = 48
    acc = np.zeros((H, W, 3))
    dn = am = None
    for f in range(frames):
        cams = hs.camera(f, W, H)
        dn, am, mv, _ = osc.gbuffer(cams[0])
        d, _, _ = osc.pt_plain(cams[0], dn, am, 2)
        acc += d[..., :3]
    est = acc / frames
    cam = hs.camera(0, W, H)[0]
    albedo = round(0.5 * 255) / 255.0  # packUnorm4x8 quantisation of the G-buffer albedo
    corners = [np.array(c, float) for c in ([-0.5, -0.5, 1.0], [0.5, -0.5, 1.0], [0.5, 0.5, 1.0], [-0.5, 0.5, 1.0])]
    ys, xs = np.mgrid[0:H, 0:W]
    want = np.zeros((H, W))
    dirs = np.zeros((H, W, 3), np.float32)
    orc.lib().orc_camera_rays(ctypes.byref(cam), dirs.ctypes.data)
    for y in range(H):
        for x in range(W):
            if dn[y, x, 0] == 0:
                continue
            p = np.array(list(cam.pos)) + dirs[y, x] * (dn[y, x, 0] - 1e-4)
            want[y, x] = albedo / np.pi * irradiance_from_quad(p, corners, 4.0, np.array([0, 0, -1.0]))
    m = dn[..., 0] > 0
    assert m.mean() > 0.9
    rel = est[..., 0][m].mean() / want[m].mean() - 1.0
    assert abs(rel) < 0.02, f"image-mean direct radiance off by {rel:.4f}"
    assert np.allclose(est[..., 0], est[..., 1]) and np.allclose(est[..., 1], est[..., 2])  # grey scene


def test_gris_final_shading_reconstructs_the_sample():
    """With both reuses off, spatialReuse's replay + reconnection (gris_resample_spatial.glsl:88-131)
    must rebuild the stream-RIS contribution F of the pixel's own sample: a check of the replay's
    random-number alignment and of the reconnection shading."""
    hs = rt.HostScene(0, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 64, 64
    n = W * H
    s = rt.settings(temporal=False, spatial=False, max_depth=6)
    cams = hs.camera(2, W, H)
    dn, am, mv, _ = osc.gbuffer(cams[0])
    resv = orc.reset_reservoirs(n)
    osc.gris_path_trace(cams, dn, am, s, resv)
    temp = orc.reset_reservoirs(n)
    osc.gris_temporal(cams, dn, am, mv, dn, am, s, resv, orc.reset_reservoirs(n), temp)
    # temporal reuse off: Temp is this frame's reservoir
    for f in orc.RESV_FIELDS:
        assert np.array_equal(np.ascontiguousarray(temp[f]).view(np.uint8), np.ascontiguousarray(resv[f]).view(np.uint8))
    out = np.zeros((H, W, 4), np.float32)
    final = orc.reset_reservoirs(n)
    osc.gris_spatial(cams, dn, am, s, temp, final, out)
    F = resv["F"].reshape(H, W, 3).astype(np.float64)
    valid = (resv["instanceIdx"] != 0xFFFFFFFF).reshape(H, W) & (F.max(-1) > 1e-6)
    assert valid.mean() > 0.3
    got = out[..., :3].astype(np.float64)
    want = np.clip(F, 0, 1e4)
    close = np.isclose(got, want, rtol=1e-3, atol=1e-6).all(-1)
    assert close[valid].mean() > 0.999, f"{int((~close[valid]).sum())} of {int(valid.sum())} samples not rebuilt"


def test_reservoir_cap_and_clear_flag():
    """M-cap (gris_reservoir.glsl:131-136) bounds sampleCount; the camera clear flag
    (gris_resample_temporal.glsl:73) disables temporal reuse."""
    hs = rt.HostScene(0, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 48, 48
    n = W * H
    cap = 4
    s = rt.settings(temporal=True, spatial=True, cap=cap, max_depth=5)
    this = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    temp = orc.reset_reservoirs(n)
    out = np.zeros((H, W, 4), np.float32)
    prev_dn = prev_am = None
    cur = 0
    for f in range(6):
        cams = hs.camera(f, W, H)
        dn, am, mv, _ = osc.gbuffer(cams[0])
        pdn = prev_dn if prev_dn is not None else np.zeros_like(dn)
        pam = prev_am if prev_am is not None else np.zeros_like(am)
        osc.gris_path_trace(cams, dn, am, s, this[cur])
        before = this[cur].copy()
        osc.gris_temporal(cams, dn, am, mv, pdn, pam, s, this[cur], this[cur ^ 1], temp)
        if f == 0:  # clear flag: no temporal merge
            for k in orc.RESV_FIELDS:
                a, b = temp[k], before[k]
                assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
        osc.gris_spatial(cams, dn, am, s, temp, this[cur], out)
        sc = this[cur]["sampleCount"]
        assert sc.max() <= cap and sc.min() >= 0
        if f >= 2:
            assert (sc > 1).mean() > 0.2  # reuse merges samples
        assert np.isfinite(out).all() and (out[..., :3] >= 0).all()
        prev_dn, prev_am = dn, am
        cur ^= 1


@pytest.mark.parametrize("unbiased", [False, True])
def test_self_reuse_weights(unbiased):
    """Spatial reuse with a zero radius merges each pixel's own reservoir again (the identity shift).
    With the unbiased weights the shaded image equals the no-reuse image pixel by pixel (the two
    copies of the sample are averaged over Z = 2, or a copy whose shift fails is not counted); the
    reference's 1/M weights halve the pixels whose identity shift fails (darkening)."""
    hs = rt.HostScene(0, 1)
    osc = orc.OracleScene(hs.arrays)
    W, H = 64, 48
    n = W * H
    base = hs.camera(0, W, H)[0]
    cams = (rt.Camera * 2)()
    ctypes.memmove(ctypes.addressof(cams), ctypes.addressof(base), ctypes.sizeof(rt.Camera))
    cams[0].frameIndex = 0x80000000
    ctypes.memmove(ctypes.addressof(cams) + ctypes.sizeof(rt.Camera), ctypes.addressof(cams), ctypes.sizeof(rt.Camera))
    dn, am, mv, _ = osc.gbuffer(cams[0])

    def render(s):
        r = orc.reset_reservoirs(n)
        t = orc.reset_reservoirs(n)
        out = np.zeros((H, W, 4), np.float32)
        osc.gris_path_trace(cams, dn, am, s, r)
        osc.gris_temporal(cams, dn, am, mv, dn, am, s, r, orc.reset_reservoirs(n), t)
        osc.gris_spatial(cams, dn, am, s, t, r, out)
        return out[..., :3].astype(np.float64), r

    none, _ = render(rt.settings(shift=2, max_depth=4, temporal=False, spatial=False, unbiased=unbiased))
    self_, r = render(rt.settings(shift=2, max_depth=4, temporal=False, spatial=True, radius=0.0, neighbors=3,
                                  unbiased=unbiased))
    lit = none.sum(-1) > 1e-3
    assert lit.mean() > 0.3
    ratio = self_.sum(-1)[lit] / none.sum(-1)[lit]
    assert (r["sampleCount"][dn.reshape(n, 4)[:, 0] > 0] == 4).all()
    if unbiased:
        assert np.abs(ratio - 1.0).max() < 1e-3
    else:
        halved = np.abs(ratio - 0.25) < 1e-3  # all three self-shifts failed: 1/4 of the weight survives
        assert np.abs(ratio[~halved] - 1.0).max() < 1e-3
        assert halved.any()
