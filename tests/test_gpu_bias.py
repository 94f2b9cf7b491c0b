"""Bias of the reuse, measured over 1024 seeds on the GPU (whose passes are bit-exact with the
oracle in every mode, test_gpu_parity.py / test_gpu_learned.py).

Each seed renders two frames of a still view from reset reservoirs; the statistic is the image
mean of the second frame's shaded output. Reuse is compared with the same GRIS estimator without
reuse (unbiased per frame: stream RIS over one path's candidates), paired by seed (both see the same
path-trace random numbers), through z = mean(d) / stderr(d) of the per-seed differences d.

* The unbiased reuse mode (1/Z contribution weights, shifted samples kept consistent with their
  new reconnection pdf, light-vertex MIS weights re-evaluated at the new predecessor) shows no bias
  with temporal + spatial reuse, for the luminance and the learned target (|z| < 4, |bias| < 2 %).
* The reference's own weights (1/M, gris_resample_spatial.glsl:128) darken the image wherever a
  shift fails; the test detects that bias (z < -4) and checks the unbiased mode removes most of it.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402

rt = gpuhelp.rt
SEEDS = 1024


def still(base, i, seed):
    c = (rt.Camera * 2)()
    C.memmove(C.addressof(c), C.addressof(base), C.sizeof(rt.Camera))
    h = (seed * 0x9E3779B1 + i * 0x85EBCA77 + 0x165667B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & 0xFFFFFFFF
    c[0].seed = h ^ (h >> 12)
    c[0].frameIndex = 0x80000000 if i == 0 else 0
    for k in range(16):
        c[0].lastProjView[k] = c[0].projView[k]
    C.memmove(C.addressof(c) + C.sizeof(rt.Camera), C.addressof(c), C.sizeof(rt.Camera))
    return c


class View:
    def __init__(self, config, W, H):
        self.hs = rt.HostScene(config, 1)
        self.ds = rt.DeviceScene(self.hs.arrays)
        self.W, self.H = W, H
        self.base = self.hs.camera(0, W, H)[0]
        self.run = gpuhelp.GrisRun(self.ds, W, H)
        self.lum = torch.tensor([0.299, 0.587, 0.114], dtype=torch.float64, device="cuda")

    def frame(self, cams, s, mlp=None, train=False, guide=None, train_guide=False):
        r = self.run
        r.gbuffer(cams[0])
        if guide is not None:
            r.guide_refresh(guide, cams)
        r.path_trace(cams, s)
        if mlp is not None:
            r.learned_refresh(mlp, cams)
        r.temporal(cams, s)
        r.spatial(cams, s)
        if train:
            rt.check(rt.core().rt_mlp_train_frame(mlp.handle, None, gpuhelp.stream()), "train")
        if train_guide:
            r.guide_train(guide, cams)
        r.flip()

    def image_means(self, s, mlp=None, frames=2, label="", guide=None):
        print(f"[gpu] bias {label}: {SEEDS} seeds x {frames} frames", flush=True)
        L, r = rt.core(), self.run
        out = []
        for seed in range(SEEDS):
            for b in r.resv + [r.temp]:
                rt.check(L.rt_gris_reservoir_reset(b.data_ptr(), self.W, self.H, gpuhelp.stream()), "reset")
            r.cur = 0
            for i in range(frames):
                self.frame(still(self.base, i, seed), s, mlp, guide=guide)
            img = r.indirect.view(torch.float32).view(self.H, self.W, 4)[..., :3].double() @ self.lum
            out.append(img.mean())
        return torch.stack(out).cpu().numpy()


def paired(a, b):
    d = a - b
    return d.mean() / b.mean(), d.mean() / (d.std(ddof=1) / np.sqrt(len(d)))


def bounds(ds, pad=0.1):
    i = ds.info()
    lo, hi = np.array(i.sceneLo), np.array(i.sceneHi)
    ext = np.maximum(hi - lo, 1e-3)
    return tuple(lo - pad * ext), tuple(hi + pad * ext)


@pytest.mark.parametrize("target", ["luminance", "learned"])
def test_unbiased_reuse_shows_no_bias(target):
    v = View(1, 128, 72)
    learned = target == "learned"
    kw = dict(shift=2, max_depth=6, cap=20, radius=7.5)
    mlp = None
    if learned:
        lo, hi = bounds(v.ds)
        mlp = rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=3))
        s_train = rt.settings(temporal=True, spatial=True, unbiased=True, learned=True, **kw)
        for i in range(24):  # train on this view, then keep the network fixed for the measurement
            v.frame(still(v.base, i, 99991), s_train, mlp, train=True)
    none = v.image_means(rt.settings(temporal=False, spatial=False, learned=learned, **kw), mlp, label="no reuse")
    unb = v.image_means(rt.settings(temporal=True, spatial=True, unbiased=True, learned=learned, **kw), mlp,
                        label=f"unbiased reuse, {target}")
    rel, z = paired(unb, none)
    print(f"[gpu] {target}: unbiased temporal+spatial reuse vs none: bias {rel:+.4f}, z {z:+.2f}", flush=True)
    assert abs(z) < 4.0 and abs(rel) < 0.02


def test_reference_weights_bias_is_detected_and_removed():
    v = View(0, 128, 96)
    kw = dict(shift=2, max_depth=4, cap=20, radius=10.0, neighbors=1, temporal=False, spatial=True)
    none = v.image_means(rt.settings(**dict(kw, spatial=False)), label="no reuse")
    ref = v.image_means(rt.settings(**kw), label="reference weights")
    unb = v.image_means(rt.settings(unbiased=True, **kw), label="unbiased weights")
    rel_ref, z_ref = paired(ref, none)
    rel_unb, z_unb = paired(unb, none)
    print(f"[gpu] box room spatial reuse: reference {rel_ref:+.4f} (z {z_ref:+.2f}), "
          f"unbiased {rel_unb:+.4f} (z {z_unb:+.2f})", flush=True)
    assert z_ref < -4.0
    # no detectable bias: the reconnection Jacobian with the geometric normal, surface vertices on
    # either side of their shading normal and the path tracer's own cosines and classification
    assert abs(z_unb) < 4.0 and abs(rel_unb) < 0.1 * abs(rel_ref)


def test_guided_unbiased_reuse_shows_no_bias():
    """The learned proposal (trained on the view, then fixed) with unbiased temporal + spatial reuse
    against the GRIS estimator without reuse or proposal. The two consume different random numbers,
    so the comparison is unpaired: z = difference of means over its standard error."""
    from test_gpu_guide import GuidedRun

    v = View(1, 128, 72)
    v.run = GuidedRun(v.ds, v.W, v.H)
    lo, hi = bounds(v.ds)
    guide = rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=5), guide=True)
    kw = dict(shift=2, max_depth=6, cap=20, radius=7.5)
    s_gu = rt.settings(temporal=True, spatial=True, unbiased=True, guided=True, **kw)
    for i in range(32):
        v.frame(still(v.base, i, 77777), s_gu, guide=guide, train_guide=True)
    none = v.image_means(rt.settings(temporal=False, spatial=False, **kw), label="no reuse, BSDF sampling")
    gu = v.image_means(s_gu, label="guided, unbiased reuse", guide=guide)
    d = gu.mean() - none.mean()
    z = d / np.sqrt(gu.var(ddof=1) / len(gu) + none.var(ddof=1) / len(none))
    print(f"[gpu] guided unbiased reuse vs none: bias {d / none.mean():+.4f}, z {z:+.2f}", flush=True)
    assert abs(z) < 4.0 and abs(d / none.mean()) < 0.02


def walk(hs, W, H, i, seed):
    """Frame i of the scene's camera walk (moving camera, lastProjView = the previous frame's), with
    a per-seed random seed; frameIndex 0 so the output is that frame's own estimate (clear flag on
    frame 0)."""
    c = hs.camera(i, W, H)
    h = (seed * 0x9E3779B1 + i * 0x85EBCA77 + 0x165667B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & 0xFFFFFFFF
    c[0].seed = c[1].seed = h ^ (h >> 12)
    c[0].frameIndex = 0x80000000 if i == 0 else 0
    return c


def test_unbiased_reuse_under_camera_motion():
    """Temporal reuse across the terrain's moving camera (reprojection by the G-buffer's motion
    vectors, rejection by normal, distance and material) plus spatial reuse, unbiased weights,
    against the estimator without reuse at the same camera: the fourth frame's image mean over 1024
    seeds, paired by seed."""
    v = View(1, 128, 72)
    kw = dict(shift=2, max_depth=6, cap=20, radius=7.5)
    L, r = rt.core(), v.run

    def means(s, label):
        print(f"[gpu] bias under motion, {label}: {SEEDS} seeds x 4 frames", flush=True)
        out = []
        for seed in range(SEEDS):
            for b in r.resv + [r.temp]:
                rt.check(L.rt_gris_reservoir_reset(b.data_ptr(), v.W, v.H, gpuhelp.stream()), "reset")
            r.cur = 0
            for i in range(4):
                v.frame(walk(v.hs, v.W, v.H, i, seed), s)
            img = r.indirect.view(torch.float32).view(v.H, v.W, 4)[..., :3].double() @ v.lum
            out.append(img.mean())
        return torch.stack(out).cpu().numpy()

    none = means(rt.settings(temporal=False, spatial=False, **kw), "no reuse")
    unb = means(rt.settings(temporal=True, spatial=True, unbiased=True, **kw), "unbiased reuse")
    rel, z = paired(unb, none)
    print(f"[gpu] moving camera: unbiased temporal+spatial reuse vs none: bias {rel:+.4f}, z {z:+.2f}", flush=True)
    assert abs(z) < 4.0 and abs(rel) < 0.02


def test_unbiased_reuse_on_glossy_metals():
    """A metallic-heavy scene (tests/golden/scenes/ggx_room.xml: GGX walls, a rough metallic box, a
    delta metal box, a glossy metal ball too smooth to reconnect on) with spatial reuse in the
    unbiased mode against no reuse, 1024 paired seeds."""
    from pathlib import Path

    import rtpkg

    hs = rt.HostScene.load(str(Path(rtpkg.ROOT) / "tests" / "golden" / "scenes" / "ggx_room.xml"))
    v = View.__new__(View)
    v.hs, v.ds, v.W, v.H = hs, rt.DeviceScene(hs.arrays), 96, 72
    v.base = hs.camera(0, v.W, v.H)[0]
    v.run = gpuhelp.GrisRun(v.ds, v.W, v.H)
    v.lum = torch.tensor([0.299, 0.587, 0.114], dtype=torch.float64, device="cuda")
    kw = dict(shift=2, max_depth=5, cap=20, radius=8.0)
    none = v.image_means(rt.settings(temporal=False, spatial=False, **kw), label="glossy metals, no reuse")
    unb = v.image_means(rt.settings(temporal=True, spatial=True, unbiased=True, **kw),
                        label="glossy metals, unbiased reuse")
    rel, z = paired(unb, none)
    print(f"[gpu] glossy metals: unbiased temporal+spatial reuse vs none: bias {rel:+.4f}, z {z:+.2f}", flush=True)
    assert abs(z) < 4.0 and abs(rel) < 0.02
