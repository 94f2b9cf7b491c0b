"""Pins the light table (AliasTable.h), light sampling (light_sampling.glsl), the
camera ray (camera.glsl), the G-buffer's motion vectors (GBuffer.frag) and the procedural
workloads."""
import ctypes as C

import numpy as np
import pytest

import orc
import rtpkg

rt = rtpkg.rt
L = orc.lib()


def oracle_alias(p):
    p = np.ascontiguousarray(p, np.float32)
    out = np.zeros((len(p) + 1, 2), np.uint32)
    L.orc_alias_build(p.ctypes.data, len(p), out.ctypes.data)
    return out[:, 0].view(np.float32), out[:, 1]


def product_alias(p):
    p = np.ascontiguousarray(p, np.float32)
    out = np.zeros((len(p) + 1, 2), np.uint32)
    rt.check(rt.host().rt_light_table_build(p.ctypes.data, len(p), out.ctypes.data), "alias")
    return out[:, 0].view(np.float32), out[:, 1]


@pytest.mark.parametrize("n", [1, 2, 7, 16, 64, 1000])
def test_alias_table_reconstructs_distribution(n):
    g = np.random.default_rng(n)
    p = g.lognormal(size=n).astype(np.float32)
    prob, fail = oracle_alias(p)
    assert prob[0] == np.float32(p.astype(np.float32).sum(dtype=np.float32)) or np.isclose(prob[0], p.sum(), rtol=1e-6)
    assert fail[0] == n
    # P(i) = (prob_i + sum_{j: fail_j = i} (1 - prob_j)) / n  (1-based buckets)
    rec = np.zeros(n + 1)
    rec[1:] += np.minimum(prob[1:], 1.0)
    for j in range(1, n + 1):
        if prob[j] < 1.0:
            rec[fail[j]] += 1.0 - prob[j]
    assert np.allclose(rec[1:] / n, p / p.sum(), atol=2e-5)
    # the product's host build is bit-identical
    pp, pf = product_alias(p)
    assert np.array_equal(prob.view(np.uint32), pp.view(np.uint32)) and np.array_equal(fail, pf)


def test_light_sampling_by_power_frequencies():
    hs = rt.HostScene(1, 1)
    lights = hs.numpy()["lights"]
    n = len(lights)
    g = np.random.default_rng(3)
    N = 200000
    q = np.zeros((N, 7), np.float32)
    q[:, :3] = [0.0, 0.0, -50.0]
    q[:, 3:] = g.random((N, 4))
    out = np.zeros((N, 11), np.float32)
    ids = np.zeros(N, np.uint32)
    L.orc_sample_light(C.byref(hs.arrays), q.ctypes.data, N, out.ctypes.data, ids.ctypes.data)
    rad = np.array([list(l["radiance"]) for l in lights])
    area = lights["area"].astype(np.float64)
    power = (rad * np.array([0.299, 0.587, 0.114])).sum(1) * area
    want = power / power.sum()
    got = np.bincount(ids, minlength=n) / N
    assert np.allclose(got, want, atol=4 * np.sqrt(want / N) + 1e-4)
    # sampled points lie on their triangle; pdf = dist^2 / (|cos| A) * (p_i)
    wi, dist, pdf = out[:, :3], out[:, 3], out[:, 4]
    v0 = np.array([list(l["v0"]) for l in lights])[ids]
    v1 = np.array([list(l["v1"]) for l in lights])[ids]
    v2 = np.array([list(l["v2"]) for l in lights])[ids]
    b = out[:, 6:8]
    pnt = v0 * (1 - b[:, :1] - b[:, 1:]) + v1 * b[:, :1] + v2 * b[:, 1:]
    ref = q[:, :3] + wi * dist[:, None]
    assert np.allclose(ref, pnt, atol=2e-4)
    nrm = np.array([[l["nx"], l["ny"], l["nz"]] for l in lights])[ids]
    cos = np.abs((nrm * wi).sum(1))
    want_pdf = dist.astype(np.float64) ** 2 / cos / area[ids] * want[ids]
    assert np.allclose(pdf, want_pdf, rtol=1e-3)


def test_camera_rays_closed_form():
    hs = rt.HostScene(1, 1)
    W, H = 64, 36
    cams = hs.camera(3, W, H)
    cam = cams[0]
    dirs = np.zeros((H, W, 3), np.float32)
    L.orc_camera_rays(C.byref(cam), dirs.ctypes.data)
    fov = np.radians(cam.FOV) / 2
    y, x = np.mgrid[0:H, 0:W]
    u, v = (x + 0.5) / W, 1 - (y + 0.5) / H
    px = (2 * u - 1) * (W / H) * np.tan(fov)
    py = (2 * v - 1) * np.tan(fov)
    R, U, F = (np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "front"))
    d = px[..., None] * R + py[..., None] * U + F
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    assert np.abs(dirs - d).max() < 2e-6
    # projView maps the centre ray to NDC (0, 0) (host Camera vs the shader's ray)
    P = np.array(list(cam.projView), np.float64).reshape(4, 4).T
    p = np.array(list(cam.pos)) + 3.0 * F
    c = P @ np.append(p, 1.0)
    assert np.allclose(c[:2] / c[3], 0.0, atol=1e-5)


def pixel_rays(cam, W, H):
    fov = np.radians(cam.FOV) / 2
    y, x = np.mgrid[0:H, 0:W]
    px = (2 * (x + 0.5) / W - 1) * (W / H) * np.tan(fov)
    py = (2 * (1 - (y + 0.5) / H) - 1) * np.tan(fov)
    R, U, F = (np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "front"))
    d = px[..., None] * R + py[..., None] * U + F
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


@pytest.mark.parametrize("frame", [0, 5, 11])
def test_motion_vectors_reproject_through_last_proj_view(frame):
    """GBuffer.frag:40-53: motion = (lastProjView p).xy / w * 0.5 + 0.5 - gl_FragCoord.xy / filmSize
    for the surface point p. Here p is rebuilt from the G-buffer alone (camera position + unit ray
    direction * the stored Euclidean depth) and projected in float64 through the camera walk's
    lastProjView; with lastProjView = projView the motion is zero (the projection and the pixel
    centres agree). Tolerance: the half-float store (2^-11 relative) plus float32 depth rounding."""
    hs = rt.HostScene(1, 1)
    W, H = 64, 36
    osc = orc.OracleScene(hs.arrays)
    cams = hs.camera(frame, W, H)
    cam = cams[0]
    for still in (True, False):
        c = rt.Camera.from_buffer_copy(bytes(cam))
        if still:
            c.lastProjView[:] = list(cam.projView)
        dn, _, mv, _ = osc.gbuffer(c)
        fg = dn[..., 0] > 0
        assert fg.mean() > 0.5
        p = np.array(list(c.pos), np.float64) + pixel_rays(c, W, H) * dn[..., :1].astype(np.float64)
        LPV = np.array(list(c.lastProjView), np.float64).reshape(4, 4).T
        q = np.concatenate([p, np.ones((H, W, 1))], -1) @ LPV.T
        last = q[..., :2] / q[..., 3:4] * 0.5 + 0.5
        y, x = np.mgrid[0:H, 0:W]
        this = np.stack([(x + 0.5) / W, (y + 0.5) / H], -1)
        ref = last - this
        got = mv.view(np.float16).astype(np.float64)
        err = np.abs(got - ref)[fg]
        assert (err <= 1e-3 * np.abs(ref[fg]) + 2e-5).all(), err.max()
        assert (got[~fg] == 0).all()
        if still:
            assert np.abs(got[fg]).max() < 1e-4
        elif frame > 0:
            assert np.abs(ref[fg]).max() > 1e-3   # the camera walk moves


def test_procedural_workloads():
    box = rt.HostScene(0, 0)
    assert box.num_triangles() - box.arrays.numLights == 12 * 2 + 64 * 320
    assert box.arrays.numLights == 2
    terr = rt.HostScene(1, 1)
    assert terr.num_triangles() - terr.arrays.numLights == 512 * 512 * 4 + 200 * 320
    assert terr.arrays.numLights == 16
    lights = terr.numpy()["lights"]
    rad = np.array([list(l["radiance"]) for l in lights])
    assert (rad.max(1) <= 100.0).all() and (rad.max(1) >= 1.0).all()
    assert (lights["nz"] < 0).all(), "emitters face down"
    mats = terr.numpy()["materials"]
    assert set(np.unique(mats["type"])) == {2}
    assert (mats["roughness"] >= 0.05).all() and (mats["roughness"] <= 1.0).all()
    assert set(np.unique(mats["metallic"])) <= {0.0, 1.0}
    # deterministic
    again = rt.HostScene(1, 1)
    assert np.array_equal(again.numpy()["vertices"].view(np.uint8), terr.numpy()["vertices"].view(np.uint8))
    c1, c2 = terr.camera(5, 32, 32), again.camera(5, 32, 32)
    assert bytes(c1) == bytes(c2)
    # uPrevCamera of frame f is uCamera of frame f-1; lastProjView is its projView
    prev = terr.camera(4, 32, 32)
    assert list(c1[1].projView) == list(prev[0].projView)
    assert list(c1[0].lastProjView) == list(prev[0].projView)
    assert terr.camera(0, 32, 32)[0].frameIndex == 0x80000000
    # the full-form scene: 1024 x 1024 cells, 2000 instanced displaced icospheres (8 shapes), 64 lights
    full = rt.HostScene(5, 4)
    assert full.num_triangles() - full.arrays.numLights == 1024 * 1024 * 4 + 2000 * 320
    assert full.arrays.numLights == 64
    a = full.numpy()
    assert (a["materials"]["roughness"] >= 0.02).all()
    verts = a["vertices"]
    sphere = verts[(1025 * 1025 + 1024 * 1024):]
    assert len(sphere) == 8 * 162
    pos = np.array([list(v) for v in sphere["pos"]], np.float64)
    nrm = np.array([list(v) for v in sphere["norm"]], np.float64)
    r = np.linalg.norm(pos, axis=1)
    assert 0.65 < r.min() < 0.95 and 1.05 < r.max() < 1.35, "spheres are displaced within +-30 %"
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    assert ((nrm * pos).sum(1) > 0).all(), "normals point outwards"
    assert not np.allclose(r[:162], r[162:324]), "the 8 shapes differ"
