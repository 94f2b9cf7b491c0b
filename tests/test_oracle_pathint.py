"""Pins the oracle's multi-bounce estimators against an independent float64 numpy Monte Carlo of
the integrals they define, on a GGX scene (tests/golden/scenes/ggx_room.xml: MetallicWorkflow
walls, a rough metallic box, a smooth metal box with a delta BSDF, a glossy metal ball too
smooth to be a reconnection vertex, a one-sided quad light):

* gi_naive (the indirect output of rt_pt_plain, gi_naive.glsl:28-149): next-event estimation at
  every secondary vertex with its MIS weight forced to 1 (:118), emission reached by BSDF sampling
  at bounce > 1 weighted by the power heuristic against the light's density (:83-93), a clamp of
  the path's radiance to [0, 1e4] (:148);
* GRIS without reuse (gris_path_trace.glsl:45-306, shaded by the spatial pass): the stream-RIS
  estimate F_i * sum(w) / w_i with w_i = luminance(F_i) has the expectation sum_i F_i over the
  path's candidates, so the image's expectation is the sum of the contributions the path tracer
  files as candidates: NEE at a vertex in state 1 (a connectible vertex after a non-connectible
  one) as a light-sampled reconnection, NEE and emission at the reconnection vertex and after it
  when that vertex is valid (gris_path_trace.glsl:111-119,137-153,198-220,260-262), with the
  path tracer's MIS weights (NEE against the Lambert density |cos|/pi, :185-186);
The numpy tracer draws its own random numbers, intersects the world-space triangles by brute
force in float64, computes shading normals from the meshes and instance matrices, and samples
lights by an inverse-CDF pick over power (not the alias table). BSDF values, densities and
samples come from the oracle's material functions, which tests/test_oracle_bsdf.py pins against
closed forms and sampling statistics. A dropped term, a wrong weight or a wrong state rule moves
the expectation by far more than the statistical tolerance (checked by planting such mistakes in
the numpy tracer: see test_tracer_detects_planted_mistakes)."""
from pathlib import Path

import numpy as np
import pytest

import orc
import rtpkg
from test_oracle_di import seeded

rt = rtpkg.rt
SCENE = Path(rtpkg.ROOT) / "tests" / "golden" / "scenes" / "ggx_room.xml"
LUM = np.array([0.299, 0.587, 0.114])
MIN_T, MAX_T = 1e-4, 1e7
GRIS_DIST = 0.01   # GRISDistanceThreshold, gris_reservoir.glsl:11
W, H, DEPTH = 32, 24, 4
LAMBERT, METAL_WF, METAL = 1, 2, 3


def mis(f, g):
    return f * f / (f * f + g * g)


def mat_struct(m):
    return rt.Material.from_buffer_copy(np.asarray(m).tobytes())


class NumpyScene:
    """World-space triangles (lights first, then each instance's), their materials and vertex
    normals, built from the scene arrays without the oracle."""

    def __init__(self, hs):
        a = hs.numpy()
        P, N, mats, light = [], [], [], []
        lights = a["lights"]
        for l in lights:
            P.append([list(l["v0"]), list(l["v1"]), list(l["v2"])])
            n = [l["nx"], l["ny"], l["nz"]]
            N.append([n, n, n])
            mats.append(-1)
        self.num_lights = len(lights)
        self.light_radiance = np.array([list(l["radiance"]) for l in lights], np.float64)
        self.light_area = np.array([l["area"] for l in lights], np.float64)
        self.light_normal = np.array([[l["nx"], l["ny"], l["nz"]] for l in lights], np.float64)
        verts, idx, midx = a["vertices"], a["indices"], a["materialIndices"]
        for inst in a["instances"]:
            M = np.array(list(inst["transform"]), np.float64).reshape(4, 4).T
            MiT = np.array(list(inst["transformInvT"]), np.float64).reshape(4, 4).T
            off, cnt = int(inst["indexOffset"]), int(inst["indexCount"])
            for k in range(cnt // 3):
                ids = idx[off + 3 * k: off + 3 * k + 3]
                pos = np.array([list(verts[i]["pos"]) for i in ids], np.float64)
                nrm = np.array([list(verts[i]["norm"]) for i in ids], np.float64)
                P.append((M[:3, :3] @ pos.T).T + M[:3, 3])
                N.append((MiT[:3, :3] @ nrm.T).T)
                mats.append(int(midx[off // 3 + k]))
        self.P = np.array(P, np.float64)            # [T, 3, 3]
        self.N = np.array(N, np.float64)            # [T, 3, 3] (object triangles: unnormalised)
        self.mat = np.array(mats, np.int64)         # -1 for lights
        self.materials = [mat_struct(m) for m in a["materials"]]
        self.mtype = np.array([m.type for m in self.materials])
        self.rough = np.array([m.roughness for m in self.materials], np.float64)
        self.metal = np.array([m.metallic for m in self.materials], np.float64)
        self.base = np.array([list(m.baseColor) for m in self.materials], np.float64)
        power = self.light_radiance @ LUM * self.light_area
        self.sum_power = power.sum()
        self.light_cdf = np.cumsum(power) / self.sum_power
        self.light_p = power / self.sum_power

    def connectible(self, m):
        t, r, mt = self.mtype[m], self.rough[m], self.metal[m]
        return np.where(t == LAMBERT, True, np.where(t == METAL_WF, (r > 0.05) | (mt < 0.9),
                                                      np.where(t == METAL, r > 0.05, False)))

    def delta(self, m):
        t, r, mt = self.mtype[m], self.rough[m], self.metal[m]
        return np.where(t == LAMBERT, False, np.where(t == METAL_WF, (r < 0.01) & (mt > 0.9),
                                                       np.where(t == METAL, r < 0.01, True)))

    def intersect(self, o, d, tmax=None):
        """Closest hit in [MIN_T, tmax): (t, triangle, b1, b2), triangle -1 on a miss
        (Moller-Trumbore in float64)."""
        n = len(o)
        best = np.full(n, np.inf if tmax is None else 0.0)
        if tmax is None:
            best[:] = MAX_T
        else:
            best = tmax.copy()
        tri = np.full(n, -1, np.int64)
        B1 = np.zeros(n)
        B2 = np.zeros(n)
        for k in range(len(self.P)):
            v0, v1, v2 = self.P[k]
            e1, e2 = v1 - v0, v2 - v0
            pv = np.cross(d, e2)
            det = pv @ e1
            ok = np.abs(det) > 1e-300
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            tv = o - v0
            b1 = (tv * pv).sum(1) * inv
            qv = np.cross(tv, e1)
            b2 = (d * qv).sum(1) * inv
            t = (qv @ e2) * inv
            hit = ok & (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1) & (t >= MIN_T) & (t < best)
            best = np.where(hit, t, best)
            tri = np.where(hit, k, tri)
            B1 = np.where(hit, b1, B1)
            B2 = np.where(hit, b2, B2)
        return best, tri, B1, B2

    def occluded(self, o, d, tmax):
        _, tri, _, _ = self.intersect(o, d, tmax)
        return tri >= 0

    def surface(self, o, d, t, tri, b1, b2):
        pos = o + d * t[:, None]
        b0 = 1.0 - b1 - b2
        nrm = self.N[tri, 0] * b0[:, None] + self.N[tri, 1] * b1[:, None] + self.N[tri, 2] * b2[:, None]
        light = tri < self.num_lights
        nrm = np.where(light[:, None], self.N[tri, 0], nrm / np.linalg.norm(nrm, axis=1, keepdims=True))
        m = np.where(light, 0, self.mat[tri])
        albedo = np.where(light[:, None], self.light_radiance[np.minimum(tri, self.num_lights - 1)], self.base[m])
        return pos, nrm, albedo, m, light

    def sample_light(self, ref, g):
        n = len(ref)
        i = np.minimum(np.searchsorted(self.light_cdf, g.random(n), side="right"), self.num_lights - 1)
        u1, u2 = g.random(n), g.random(n)
        s = np.sqrt(u1)
        b1, b2 = 1 - s, u2 * s
        v = self.P[i]
        pos = v[:, 0] * (1 - b1 - b2)[:, None] + v[:, 1] * b1[:, None] + v[:, 2] * b2[:, None]
        diff = pos - ref
        dist = np.linalg.norm(diff, axis=1)
        wi = diff / dist[:, None]
        nl = self.light_normal[i]
        cosl = (nl * wi).sum(1)
        pdf = dist ** 2 / (np.abs(cosl) * self.light_area[i]) * self.light_p[i]
        Le = np.where((cosl > 0)[:, None], 0.0, self.light_radiance[i])
        return Le, wi, dist, pdf

    def bsdf_eval(self, m, albedo, n, wo, wi):
        f = np.zeros((len(m), 3))
        for k in np.unique(m):
            sel = m == k
            q = np.ascontiguousarray(np.concatenate([n[sel], wo[sel], wi[sel]], 1).astype(np.float32))
            f[sel] = self._eval_rows(k, albedo[sel].astype(np.float32), q)
        return f

    def _eval_rows(self, k, alb, q):
        # the oracle evaluates one albedo per call: group rows by albedo
        out = np.zeros((len(q), 3))
        keys, inv = np.unique(alb, axis=0, return_inverse=True)
        for a_i, a in enumerate(keys):
            rows = np.nonzero(inv.reshape(-1) == a_i)[0]
            qq = np.ascontiguousarray(q[rows])
            fk = np.zeros((len(rows), 3), np.float32)
            pk = np.zeros(len(rows), np.float32)
            aa = np.ascontiguousarray(a, np.float32)
            orc.lib().orc_eval_bsdf(self.materials[k], aa.ctypes.data, qq.ctypes.data, len(rows), fk.ctypes.data,
                                    pk.ctypes.data)
            out[rows] = fk
        return out

    def bsdf_sample(self, m, albedo, n, wo, g):
        k_all = len(m)
        wi = np.zeros((k_all, 3))
        f = np.zeros((k_all, 3))
        pdf = np.zeros(k_all)
        typ = np.zeros(k_all, np.int64)
        ok = np.zeros(k_all, bool)
        r = g.random((k_all, 3))
        for k in np.unique(m):
            sel = np.nonzero(m == k)[0]
            keys, inv = np.unique(albedo[sel].astype(np.float32), axis=0, return_inverse=True)
            for a_i, a in enumerate(keys):
                rows = sel[inv.reshape(-1) == a_i]
                q = np.ascontiguousarray(np.concatenate([n[rows], wo[rows], r[rows]], 1).astype(np.float32))
                c = len(rows)
                wk = np.zeros((c, 3), np.float32)
                fk = np.zeros((c, 3), np.float32)
                pk = np.zeros(c, np.float32)
                tk = np.zeros(c, np.uint32)
                ok_k = np.zeros(c, np.uint32)
                aa = np.ascontiguousarray(a, np.float32)
                orc.lib().orc_sample_bsdf(self.materials[k], aa.ctypes.data, q.ctypes.data, c, wk.ctypes.data,
                                          fk.ctypes.data, pk.ctypes.data, tk.ctypes.data, ok_k.ctypes.data)
                wi[rows], f[rows], pdf[rows], typ[rows], ok[rows] = wk, fk, pk, tk, ok_k != 0
        return wi, f, pdf, typ, ok


def trace(sc, o, d, depth, n0, alb0, mat0, g, mode, plant=None):
    """Per-path estimates of the gi_naive ('gi') or GRIS-candidate-sum ('gris') integrands.
    o, d: camera rays; depth: G-buffer depth; n0, alb0, mat0: G-buffer normal, albedo, material.
    plant: a deliberate mistake for the self-test ('nee_mis' uses the MIS weight in gi_naive,
    'no_state' ignores the reconnection-vertex rules, 'light_mis' drops the emission MIS weight)."""
    N = len(o)
    out = np.zeros((N, 3))
    pos = o + d * (depth - 1e-4)[:, None]
    nrm, alb, m = n0.copy(), alb0.copy(), mat0.copy()
    wo = -d
    thr = np.ones((N, 3))
    alive = np.ones(N, bool)
    last_pos = np.zeros((N, 3))
    dirv = d.copy()
    s_pdf = np.zeros(N)
    s_delta = np.zeros(N, bool)
    # GRIS bookkeeping
    state = np.zeros(N, np.int64)
    last_conn = np.zeros(N, bool)
    rc_valid = np.zeros(N, bool)     # the path's reconnection vertex was recorded
    acc = np.zeros((N, 3))           # contributions filed under the reconnection vertex (state 2, 2)
    light_hit = np.zeros(N, bool)
    for bounce in range(depth_limit(depth)):
        if bounce > 0:
            idx = np.nonzero(alive)[0]
            t, tri, b1, b2 = sc.intersect(pos[idx] + dirv[idx] * 1e-4, dirv[idx])
            miss = tri < 0
            alive[idx[miss]] = False
            idx, t, tri, b1, b2 = idx[~miss], t[~miss], tri[~miss], b1[~miss], b2[~miss]
            last_pos[idx] = pos[idx]
            p2, n2, a2, m2, l2 = sc.surface(pos[idx] + dirv[idx] * 1e-4, dirv[idx], t, tri, b1, b2)
            pos[idx], nrm[idx], alb[idx], m[idx] = p2, n2, a2, m2
            light_hit[:] = False
            light_hit[idx] = l2
        idx = np.nonzero(alive)[0]
        if len(idx) == 0:
            break
        is_light = light_hit[idx] if bounce > 0 else np.zeros(len(idx), bool)
        conn_this = np.where(is_light, True, sc.connectible(m[idx]))
        dist_prev = np.linalg.norm(last_pos[idx] - pos[idx], axis=1)
        last_state = state[idx].copy()
        st = np.where(last_state == 2, 2, np.where(conn_this, last_state + 1, 0))
        state[idx] = st
        # emission reached by the BSDF sample of the previous vertex
        li = idx[is_light]
        if len(li):
            cos_prev = (dirv[li] * nrm[li]).sum(1)
            front = cos_prev < 0
            ok = front & (bounce > 1)
            dist2 = ((last_pos[li] - pos[li]) ** 2).sum(1)
            light_pdf = (alb[li] @ LUM) / sc.sum_power * dist2 / np.abs(cos_prev)
            w = np.where(s_delta[li], 1.0, mis(s_pdf[li], light_pdf))
            if plant == "light_mis":
                w = np.ones_like(w)
            c = alb[li] * w[:, None] * thr[li]
            c[~ok] = 0.0
            if mode == "gi":
                out[li] += c
            else:
                sl, ls = st[is_light], last_state[is_light]
                both2 = (sl == 2) & (ls == 2)
                first = (sl == 2) & (ls == 1) & last_conn[li] & (dist_prev[is_light] > GRIS_DIST)
                if plant == "no_state":
                    out[li] += c
                else:
                    acc[li] += np.where((both2 & rc_valid[li])[:, None], c, 0.0)
                    out[li] += np.where(first[:, None], c, 0.0)
            alive[li] = False
        idx = idx[~is_light]
        st, last_state = st[~is_light], last_state[~is_light]
        conn_this, dist_prev = conn_this[~is_light], dist_prev[~is_light]
        if len(idx) == 0:
            break
        if mode == "gris":
            # the reconnection vertex: first connectible vertex after a connectible one, far enough
            becomes = (st == 2) & (last_state == 1)
            valid = becomes & conn_this & last_conn[idx] & (dist_prev > GRIS_DIST)
            rc_valid[idx[becomes]] = valid[becomes]
        # next-event estimation at secondary vertices with a non-delta BSDF
        nee = idx[(bounce > 0) & ~sc.delta(m[idx])]
        if len(nee):
            Le, wi, dist, lpdf = sc.sample_light(pos[nee], g)
            shadowed = sc.occluded(pos[nee], wi, dist - 1e-4)
            ok = ~shadowed & (lpdf > 1e-6)
            f = sc.bsdf_eval(m[nee], alb[nee], nrm[nee], wo[nee], wi)
            cosn = np.maximum((nrm[nee] * wi).sum(1), 0.0)
            if mode == "gi" and plant != "nee_mis":
                w = np.ones(len(nee))
            else:
                w = mis(lpdf, np.abs((nrm[nee] * wi).sum(1)) / np.pi)
            c = Le / lpdf[:, None] * w[:, None] * f * cosn[:, None] * thr[nee]
            c[~ok] = 0.0
            if mode == "gi":
                out[nee] += c
            else:
                sel = np.isin(idx, nee)
                st_n, ls_n, conn_n = st[sel], last_state[sel], conn_this[sel]
                both2 = (st_n == 2) & (ls_n == 2)
                first = (st_n == 2) & (ls_n == 1)
                light_rc = (st_n == 1) & conn_n & (dist > GRIS_DIST)
                if plant == "no_state":
                    out[nee] += c
                else:
                    acc[nee] += np.where((both2 & rc_valid[nee])[:, None], c, 0.0)
                    out[nee] += np.where(((first & rc_valid[nee]) | light_rc)[:, None], c, 0.0)
        # continuation by BSDF sampling
        wi, f, pdf, typ, ok = sc.bsdf_sample(m[idx], alb[idx], nrm[idx], wo[idx], g)
        ok &= pdf >= 1e-6
        dead = idx[~ok]
        alive[dead] = False
        idx, wi, f, pdf, typ = idx[ok], wi[ok], f[ok], pdf[ok], typ[ok]
        is_delta = (typ & 4) == 4
        cos = np.where(is_delta, 1.0, np.abs((nrm[idx] * wi).sum(1)))
        thr[idx] *= f * (cos / pdf)[:, None]
        s_pdf[idx], s_delta[idx] = pdf, is_delta
        last_conn[idx] = conn_this[ok]
        wo[idx] = -wi
        dirv[idx] = wi
    if mode == "gris":
        out += acc
    else:
        out = np.where(np.isnan(out).any(1, keepdims=True), 0.0, np.clip(out, 0.0, 1e4))
    return out


def depth_limit(depth):
    return DEPTH


def camera_rays(cam):
    """Camera ray directions at pixel centres from the camera basis (the closed form that
    tests/test_oracle_scene.py pins), float64."""
    fov = np.radians(cam.FOV) / 2
    y, x = np.mgrid[0:H, 0:W]
    u, v = (x + 0.5) / W, 1 - (y + 0.5) / H
    px = (2 * u - 1) * (W / H) * np.tan(fov)
    py = (2 * v - 1) * np.tan(fov)
    R, U, F = (np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "front"))
    d = px[..., None] * R + py[..., None] * U + F
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


@pytest.fixture(scope="module")
def setup():
    hs = rt.HostScene.load(str(SCENE))
    osc = orc.OracleScene(hs.arrays)
    cam = hs.camera(0, W, H)[0]
    dn, am, mv, _ = osc.gbuffer(cam)
    return hs, osc, cam, dn, am, mv


def numpy_image(sc, cam, dn, am, spp, seed, mode, plant=None):
    """spp numpy paths per foreground pixel: (foreground mask, estimates [pixels, spp, 3])."""
    fg = dn[..., 0].reshape(-1) > 0
    npx = int(fg.sum())
    rep = lambda a: np.repeat(a, spp, axis=0)
    d = rep(camera_rays(cam).reshape(-1, 3)[fg])
    o = np.tile(np.array(list(cam.pos), np.float64), (len(d), 1))
    depth = rep(dn[..., 0].reshape(-1)[fg].astype(np.float64))
    n0 = rep(dn[..., 1:4].reshape(-1, 3)[fg].astype(np.float64))
    packed = rep(am[..., 0].reshape(-1)[fg])
    alb0 = np.stack([(packed >> (8 * k)) & 0xFF for k in range(3)], 1) / 255.0
    mat0 = rep((am[..., 1].reshape(-1)[fg] >> 16).astype(np.int64))
    g = np.random.default_rng(seed)
    est = trace(sc, o, d, depth, n0, alb0, mat0, g, mode, plant)
    return fg, est.reshape(npx, spp, 3)


def oracle_frames(osc, cam, dn, am, mv, frames, mode):
    n = W * H
    imgs = np.zeros((frames, n, 3))
    s = rt.settings(shift=2, temporal=False, spatial=False, max_depth=DEPTH)
    for f in range(frames):
        cams = seeded(cam, f)
        if mode == "gi":
            _, ind, _ = osc.pt_plain(cams[0], dn, am, DEPTH)
            imgs[f] = ind[..., :3].reshape(n, 3)
        else:
            r, t = orc.reset_reservoirs(n), orc.reset_reservoirs(n)
            out = np.zeros((H, W, 4), np.float32)
            osc.gris_path_trace(cams, dn, am, s, r)
            osc.gris_temporal(cams, dn, am, mv, dn, am, s, r, orc.reset_reservoirs(n), t)
            osc.gris_spatial(cams, dn, am, s, t, r, out)
            imgs[f] = out[..., :3].reshape(n, 3)
    return imgs


def region_z(fg, ora, nump, regions=4):
    """z-scores of the region means (image cut into vertical bands, per channel) of the oracle's
    frames against the numpy samples."""
    cols = np.tile(np.arange(W), H)[fg]
    band = cols * regions // W
    zs = []
    o = ora[:, fg]           # frames x pixels x 3
    for b in range(regions):
        sel = band == b
        a = o[:, sel].mean(1)                    # per frame: region mean (frames x 3)
        c = nump[sel].mean(0)                    # per numpy sample: region mean (spp x 3)
        se = np.sqrt(a.var(0, ddof=1) / len(a) + c.var(0, ddof=1) / len(c))
        zs.append((a.mean(0) - c.mean(0)) / se)
    return np.array(zs), o.mean(), nump.mean()


FRAMES, SPP = 160, 160


@pytest.mark.parametrize("mode", ["gi", "gris"])
def test_oracle_matches_numpy_path_integral(setup, mode):
    hs, osc, cam, dn, am, mv = setup
    sc = NumpyScene(hs)
    ora = oracle_frames(osc, cam, dn, am, mv, FRAMES, mode)
    fg, nump = numpy_image(sc, cam, dn, am, SPP, seed=11 if mode == "gi" else 12, mode=mode)
    z, mo, mn = region_z(fg, ora, nump)
    print(f"{mode}: oracle mean {mo:.5f} numpy mean {mn:.5f} z {np.round(z, 2).tolist()}")
    assert mo > 0.01
    assert np.abs(z).max() < 4.5, z


@pytest.mark.parametrize("mode,plant", [("gi", "nee_mis"), ("gris", "no_state"), ("gris", "light_mis")])
def test_tracer_detects_planted_mistakes(setup, mode, plant):
    """The comparison has the power to see a plausible mistake: each planted one moves the numpy
    expectation well outside the tolerance the parity test uses."""
    hs, osc, cam, dn, am, mv = setup
    sc = NumpyScene(hs)
    ora = oracle_frames(osc, cam, dn, am, mv, FRAMES, mode)
    fg, nump = numpy_image(sc, cam, dn, am, SPP, seed=13, mode=mode, plant=plant)
    z, _, _ = region_z(fg, ora, nump)
    assert np.abs(z).max() > 6.0, z
