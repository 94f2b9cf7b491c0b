"""The learned directional proposal on the GPU (no counterpart in the reference), through the C ABI.

* The proposal network (32 outputs: an 8-lobe vMF mixture): MFMA inference and loss gradients
  against the oracle (FP16 rounding points mirrored, tolerances as for the radiance network);
  online training fits a synthetic direction field.
* rt_gris_guide_refresh writes the lobe planes the oracle derives from the network's outputs at
  the same primary vertices.
* GRIS with the proposal on: the oracle takes the GPU's lobe planes as input and restates the
  rest; the path tracer's reservoirs and training records, the temporal and spatial passes are
  bit-exact over frame sequences with camera motion and training between frames, with the
  reference's weights and in the unbiased mode.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402
from test_gpu_learned import bits_equal, scene_bounds  # noqa: E402
from test_gpu_mlp import BASE, HI, LO, LOG2T, MAXR, dev  # noqa: E402
from test_gpu_parity import close  # noqa: E402

rt = gpuhelp.rt


def make_guide(seed=5, params=True):
    print(f"[gpu] guide seed {seed}", flush=True)
    m = rt.Mlp(rt.mlp_config(LO, HI, log2_table=LOG2T, base_res=BASE, max_res=MAXR, seed=seed), guide=True)
    if not params:
        return m, None, None
    g = np.random.default_rng(seed)
    grid = g.uniform(-0.5, 0.5, m.grid_params).astype(np.float32)
    w = np.concatenate([g.uniform(-1, 1, 2048) * np.sqrt(6 / 96), g.uniform(-1, 1, 4096) * np.sqrt(6 / 128),
                        g.uniform(-1, 1, 2048) * np.sqrt(6 / 96)]).astype(np.float32)
    m.set_params(grid, w)
    return m, grid, w


def infer32(m, pos):
    d_pos = dev(pos)  # kept alive until the kernel has run
    out = torch.zeros((len(pos), 32), dtype=torch.float32, device="cuda")
    rt.check(rt.core().rt_mlp_infer(m.handle, d_pos.data_ptr(), len(pos), out.data_ptr(), gpuhelp.stream()), "infer")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_guide_params_and_init():
    m, _, _ = make_guide(1, params=False)
    grid, w = m.params()
    assert w.shape == (8192,) and grid.shape == (16 * (1 << LOG2T) * 2,)
    assert np.abs(w[6144:]).max() <= np.sqrt(6 / 96) and np.abs(w[6144:]).max() > 0


def test_guide_inference_matches_oracle():
    m, grid, w = make_guide(5)
    g = np.random.default_rng(1)
    for n in (20000, 1, 33):
        pos = g.uniform(np.array(LO) - 0.3, np.array(HI) + 0.3, (n, 3)).astype(np.float32)
        got = infer32(m, pos)
        want = orc.OracleGuide(LOG2T, BASE, MAXR, LO, HI).forward(grid, w, pos, mode=1)
        err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
        print(f"[gpu] guide inference n={n} err median {np.median(err):.2e} max {err.max():.2e}", flush=True)
        assert np.median(err) < 1e-5 and err.max() < 2e-2


def test_guide_gradients_match_oracle():
    m, grid, w = make_guide(7)
    g = np.random.default_rng(3)
    n = 4096 + 17
    pos = g.uniform(LO, HI, (n, 3)).astype(np.float32)
    d = g.normal(0, 1, (n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tgt = np.concatenate([d, np.exp(g.normal(0, 1, (n, 1)))], 1).astype(np.float32)
    tgt[::9, 3] = -1.0
    d_pos, d_t = dev(pos), dev(tgt)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    rt.check(rt.core().rt_mlp_gradients(m.handle, d_pos.data_ptr(), d_t.data_ptr(), n, loss.data_ptr(),
                                        gpuhelp.stream()), "gradients")
    torch.cuda.synchronize()
    gg, gw = m.read_gradients()
    oloss, ogg, ogw = orc.OracleGuide(LOG2T, BASE, MAXR, LO, HI).gradients(grid, w, pos, tgt, mode=1)
    print(f"[gpu] guide loss gpu {float(loss.item()):.6f} oracle {oloss:.6f}", flush=True)
    assert abs(float(loss.item()) - oloss) <= 1e-3 * abs(oloss)
    for name, a, b in (("W1", gw[:2048], ogw[:2048]), ("W2", gw[2048:6144], ogw[2048:6144]),
                       ("W3", gw[6144:], ogw[6144:]), ("grid", gg, ogg)):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[gpu] guide {name} gradient relative error {err:.2e}", flush=True)
        assert err < 2e-2, name
    assert np.array_equal(gg != 0, ogg != 0)


def test_guide_training_fits_a_direction_field():
    """Directions drawn from a vMF (kappa 30) about an axis that varies with position: after online
    training the mixture's density at the axis is far above uniform and its log-likelihood rises."""
    m, _, _ = make_guide(9, params=False)
    g = np.random.default_rng(5)
    centre = (np.array(LO) + np.array(HI)) / 2

    def batch(n):
        pos = g.uniform(LO, HI, (n, 3))
        axis = pos - centre
        axis /= np.linalg.norm(axis, axis=1, keepdims=True)
        # vMF(axis, 30) samples by Jakob's inversion in the axis frame
        u1, u2 = g.random(n), g.random(n)
        k = 30.0
        wz = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * k)) / k
        r = np.sqrt(np.maximum(0, 1 - wz ** 2))
        t = np.where(np.abs(axis[:, 2:3]) > 0.999, [[0, 1, 0]], [[0, 0, 1]])
        b = np.cross(axis, t)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        t = np.cross(b, axis)
        d = t * (r * np.cos(2 * np.pi * u2))[:, None] + b * (r * np.sin(2 * np.pi * u2))[:, None] + axis * wz[:, None]
        return pos.astype(np.float32), np.concatenate([d, np.ones((n, 1))], 1).astype(np.float32), axis

    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    losses = []
    for step in range(400):
        pos, tgt, _ = batch(8192)
        d_pos, d_tgt = dev(pos), dev(tgt)
        rt.check(rt.core().rt_mlp_train_step(m.handle, d_pos.data_ptr(), d_tgt.data_ptr(), len(pos), loss.data_ptr(),
                                             gpuhelp.stream()), "train")
        if step % 50 == 0 or step == 399:
            torch.cuda.synchronize()
            losses.append(float(loss.item()))
    print(f"[gpu] guide training losses {[round(x, 3) for x in losses]}", flush=True)
    pos, _, axis = batch(2000)
    lobes = orc.guide_lobes(infer32(m, pos))
    dens = np.array([orc.mixture(lobes[i], axis[i:i + 1].astype(np.float32), np.zeros((0, 3)))[0][0]
                     for i in range(len(pos))])
    print(f"[gpu] guide density at the axis: median {np.median(dens):.2f} (uniform {1 / (4 * np.pi):.3f})", flush=True)
    assert losses[-1] < losses[0] - 1.0
    assert np.median(dens) > 20 / (4 * np.pi)


def blocks(W, H):
    """The mixture is evaluated once per 2x2 pixel block."""
    return ((W + 1) // 2) * ((H + 1) // 2)


class GuidedRun(gpuhelp.GrisRun):
    """GrisRun plus the proposal's bindings: two lobe-plane parities and the training records."""

    def __init__(self, ds, W, H):
        super().__init__(ds, W, H)
        lb = rt.core().rt_guide_lobe_bytes(W, H)
        self.lobes = [gpuhelp.dev_empty(lb), gpuhelp.dev_empty(lb)]
        self.samples = gpuhelp.dev_empty(16 * W * H)

    def frame_struct(self):
        f = super().frame_struct()
        f.guideLobes = self.lobes[self.cur].data_ptr()
        f.guideLobesPrev = self.lobes[self.cur ^ 1].data_ptr()
        f.guideSamples = self.samples.data_ptr()
        return f

    def guide_refresh(self, guide, cams):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_guide_refresh(self.ds.handle, guide.handle, cams, C.byref(f), gpuhelp.stream()),
                 "guide refresh")

    def guide_train(self, guide, cams):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_guide_train(self.ds.handle, guide.handle, cams, C.byref(f), None, gpuhelp.stream()),
                 "guide train")

    def lobes_host(self, parity):
        return gpuhelp.to_host(self.lobes[parity], np.float32, (16, blocks(self.W, self.H), 4))

    def samples_host(self):
        return gpuhelp.to_host(self.samples, np.float32, (self.W * self.H, 4))


def test_guide_refresh_planes_match_oracle():
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 96, 64
    lo, hi = scene_bounds(ds)
    guide = rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=3), guide=True)
    g = np.random.default_rng(13)  # parameters with outputs of order 1: lobes of every shape
    grid0 = g.uniform(-0.5, 0.5, guide.grid_params).astype(np.float32)
    w0 = np.concatenate([g.uniform(-1, 1, 2048) * np.sqrt(6 / 96), g.uniform(-1, 1, 4096) * np.sqrt(6 / 128),
                         g.uniform(-1, 1, 2048) * np.sqrt(6 / 96)]).astype(np.float32)
    guide.set_params(grid0, w0)
    run = GuidedRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    run.guide_refresh(guide, cams)
    s = rt.settings(shift=2, max_depth=6, guided=True)
    run.path_trace(cams, s)
    run.guide_train(guide, cams)  # gathers the batch (primary positions) before its training step
    pos, tgt = guide.batch()
    lobes = run.lobes_host(run.cur)
    n = W * H
    stride = max(1, (n + 16383) // 16384)  # 2^14 training samples per frame
    offset = cams[0].seed % stride
    pix = offset + stride * np.arange(len(pos))
    dn, _, _ = run.gbuffer_host()
    fg = dn.reshape(-1, 4)[pix, 0] > 0
    assert fg.sum() > 1000
    # the planes against the oracle's conversion of the network's outputs at each 2x2 block's first
    # foreground pixel (its primary position from the gathered batch: stride 1 here), the parameters
    # restored to those the refresh ran with
    assert stride == 1 and offset == 0
    guide.set_params(grid0, w0)
    fgpix = dn.reshape(-1, 4)[:, 0] > 0
    bw = (W + 1) // 2
    reps, bids = [], []
    for b in range(blocks(W, H)):
        bx, by = 2 * (b % bw), 2 * (b // bw)
        for c in range(4):
            x, y = bx + (c & 1), by + (c >> 1)
            if x < W and y < H and fgpix[y * W + x]:
                reps.append(y * W + x)
                bids.append(b)
                break
    reps, bids = np.array(reps), np.array(bids)
    assert len(bids) > 500
    want = orc.guide_lobes(infer32(guide, pos[reps]))
    got = lobes[:, bids, :].transpose(1, 0, 2)
    err = np.abs(got - want)
    tol = 1e-4 * np.maximum(np.abs(want), 1.0)  # FP32 device exp/sin/cos against double
    bad = (err > tol).any(axis=(1, 2))
    print(f"[gpu] refresh planes: {int(bad.sum())} of {len(bad)} blocks outside tolerance, "
          f"max err {err.max():.2e}, kappa range {want[:, 0::2, 3].min():.3g}..{want[:, 0::2, 3].max():.3g}", flush=True)
    assert bad.mean() < 1e-3
    # training records: the samples the batch holds are the path tracer's records at those pixels
    rec = run.samples_host()[pix[fg]]
    keep = (rec[:, 3] > 0) & (np.linalg.norm(rec[:, :3], axis=1) > 0.5)
    assert keep.sum() > 100  # paths with a positive contribution (many terrain paths find no light)
    assert np.array_equal(tgt[fg][keep], rec[keep]) and (tgt[fg][~keep][:, 3] == -1).all()


@pytest.mark.parametrize("config,W,H,unbiased", [(0, 80, 48, False), (1, 96, 64, False), (1, 96, 64, True)],
                         ids=["boxroom", "terrain", "terrain-unbiased"])
def test_guided_gris_parity(config, W, H, unbiased):
    print(f"[gpu] guided GRIS config {config} {W}x{H} unbiased={unbiased}", flush=True)
    hs = rt.HostScene(config, 1)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    lo, hi = scene_bounds(ds)
    guide = rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=3), guide=True)
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=6, cap=20, radius=20.0, unbiased=unbiased,
                    guided=True)
    run = GuidedRun(ds, W, H)
    n = W * H
    o_dn, o_am = [None, None], [None, None]
    o_resv = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    o_temp = orc.reset_reservoirs(n)
    o_out = np.zeros((H, W, 4), np.float32)
    nb = blocks(W, H)
    o_lobes = [np.zeros((16, nb, 4), np.float32), np.zeros((16, nb, 4), np.float32)]
    o_samples = np.zeros((n, 4), np.float32)
    cur = 0
    guided_px = 0
    for frame in range(4):
        print(f"[gpu] guided frame {frame}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        run.guide_refresh(guide, cams)
        o_lobes[cur] = run.lobes_host(run.cur)
        osc.bind_guide(o_lobes[cur], o_lobes[prev], o_samples)
        run.path_trace(cams, s)
        osc.gris_path_trace(cams, dn, am, s, o_resv[cur])
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame} path trace: {int(bad.sum())} reservoirs differ"
        fg = dn.reshape(-1, 4)[:, 0] > 0
        g_rec = run.samples_host()
        assert bits_equal(g_rec[fg], o_samples[fg]), f"frame {frame}: training records differ"
        guided_px += int((np.linalg.norm(g_rec[fg, :3], axis=1) > 0.5).sum())
        run.temporal(cams, s)
        osc.gris_temporal(cams, dn, am, mv, pdn, pam, s, o_resv[cur], o_resv[prev], o_temp)
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.temp), o_temp)
        assert not bad.any(), f"frame {frame} temporal: {int(bad.sum())} reservoirs differ"
        run.spatial(cams, s)
        osc.gris_spatial(cams, dn, am, s, o_temp, o_resv[cur], o_out)
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame} spatial: {int(bad.sum())} reservoirs differ"
        out = run.output(1)
        ok = close(out, o_out).all(axis=-1)
        assert ok.all(), f"frame {frame}: {int((~ok).sum())} output pixels differ"
        run.guide_train(guide, cams)
        run.flip()
        cur ^= 1
    osc.bind_guide()
    assert guided_px > 0, "no path sampled a direction at its primary vertex"


def test_guided_requires_bindings():
    hs = rt.HostScene(2, 0)
    ds = rt.DeviceScene(hs.arrays)
    run = gpuhelp.GrisRun(ds, 16, 16)
    cams = hs.camera(0, 16, 16)
    f = run.frame_struct()
    s = rt.settings(guided=True)
    assert rt.core().rt_gris_path_trace(ds.handle, cams, C.byref(f), C.byref(s), gpuhelp.stream()) != 0
    assert b"guideLobes" in rt.core().rt_last_error()
    s = rt.settings(guided=2)
    assert rt.core().rt_gris_temporal(ds.handle, cams, C.byref(f), C.byref(s), gpuhelp.stream()) != 0
    radiance = rt.Mlp(rt.mlp_config((0, 0, 0), (1, 1, 1), log2_table=12, seed=1))
    f.guideLobes = run.direct.data_ptr()
    assert rt.core().rt_gris_guide_refresh(ds.handle, radiance.handle, cams, C.byref(f), gpuhelp.stream()) != 0
    assert b"proposal network" in rt.core().rt_last_error()


def test_renderer_with_learned_proposal():
    """The host frame driver with the proposal enabled: guided frames of the unbiased estimator run,
    train the network and give finite, positive images; the same camera path without the proposal
    gives an image of the same mean within the noise."""
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 64, 48
    Hh = rt.host()
    means = {}
    for guided in (False, True):
        r = C.c_void_p()
        rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
        try:
            if guided:
                g = C.c_void_p()
                rt.check(Hh.rt_renderer_enable_proposal(r, 4, 14, gpuhelp.stream(), C.byref(g)), "enable proposal")
            s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=4, unbiased=True, guided=guided)
            rt.check(Hh.rt_renderer_set_settings(r, C.byref(s)), "settings")
            acc = np.zeros((H, W, 3))
            for frame in range(48):
                rt.check(Hh.rt_renderer_frame(r, hs.camera(frame, W, H), 0, gpuhelp.stream()), "frame")
                host = np.zeros((H, W, 4), np.float32)
                rt.check(Hh.rt_renderer_read_output(r, 1, host.ctypes.data, gpuhelp.stream()), "read")
                assert np.isfinite(host).all() and (host[..., :3] >= 0).all()
                acc += host[..., :3]
            means[guided] = acc.mean() / 48
            if guided:
                pos = np.zeros((1 << 16, 3), np.float32)
                n = C.c_uint32()
                rt.check(rt.core().rt_mlp_batch_download(g, pos.ctypes.data, None, C.byref(n)), "batch")
                assert n.value == W * H
        finally:
            Hh.rt_renderer_destroy(r)
    print(f"[gpu] renderer mean radiance: BSDF sampling {means[False]:.4f}, proposal {means[True]:.4f}", flush=True)
    assert means[True] > 0 and abs(means[True] - means[False]) < 0.1 * means[False]
