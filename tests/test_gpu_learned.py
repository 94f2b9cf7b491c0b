"""GRIS with the learned target function against the oracle, pass by pass.

The network runs on the GPU (its numerics are checked against the oracle in
test_gpu_mlp.py); the oracle takes the GPU's per-pixel predictions as input and
restates everything else: the refresh (rescaled resample weights, learned planes,
training batch) and the learned temporal and spatial reuse are bit-exact. The
network is trained between frames so every frame uses new predictions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt
REL_TOL = 1e-4


def scene_bounds(ds, pad=0.1):
    i = ds.info()
    lo, hi = np.array(i.sceneLo), np.array(i.sceneHi)
    ext = np.maximum(hi - lo, 1e-3)
    return tuple(lo - pad * ext), tuple(hi + pad * ext)


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("config,W,H,unbiased,guided,log2,depth",
                         [(0, 96, 64, False, False, 16, 6), (1, 160, 96, False, False, 16, 6),
                          (1, 160, 96, True, False, 16, 6), (1, 160, 96, True, True, 16, 6),
                          (1, 640, 360, False, False, 19, 6), (5, 128, 72, True, True, 19, 8)],
                         ids=["boxroom", "terrain", "terrain-unbiased", "terrain-unbiased-proposal",
                              "terrain-640x360-table2^19", "scene4.8M-depth8-full-form"])
def test_learned_gris_parity(config, W, H, unbiased, guided, log2, depth):
    """With guided: the learned vMF-mixture proposal on as well (the full form of the learned
    estimator), its lobe planes taken from the GPU, its network trained between frames.
    terrain-640x360-table2^19: the learned-target workload's network (2^19 entries per level) at a
    view of more than 2^16 pixels, so the training batch takes every stride-th pixel (stride 4)
    with a per-frame offset. scene4.8M-depth8-full-form: the full-form workload's scene (~4.8M
    triangles, 64 emissive, roughness U[0.02, 1]) at depth 8 with the learned target, the learned
    proposal and unbiased reuse, at a reduced view."""
    print(f"[gpu] learned parity config {config} {W}x{H} unbiased {unbiased} proposal {guided} table 2^{log2} "
          f"depth {depth}", flush=True)
    hs = rt.HostScene(config, 1)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    lo, hi = scene_bounds(ds)
    mlp = rt.Mlp(rt.mlp_config(lo, hi, log2_table=log2, seed=3))
    omlp = orc.OracleMlp(log2, 16, 2048, lo, hi)
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=depth, cap=20, radius=20.0, learned=True,
                    unbiased=unbiased, guided=guided)
    if guided:
        from test_gpu_guide import GuidedRun
        run = GuidedRun(ds, W, H)
        guide = rt.Mlp(rt.mlp_config(lo, hi, log2_table=log2, seed=4), guide=True)
    else:
        run = gpuhelp.GrisRun(ds, W, H)
    n = W * H
    nb = ((W + 1) // 2) * ((H + 1) // 2)  # one mixture per 2x2 pixel block
    o_lobes = [np.zeros((16, nb, 4), np.float32), np.zeros((16, nb, 4), np.float32)]
    o_samples = np.zeros((n, 4), np.float32)
    o_resv = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    o_learn = [np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)]
    o_temp = orc.reset_reservoirs(n)
    o_learn_temp = np.zeros((n, 4), np.float32)
    o_out = np.zeros((H, W, 4), np.float32)
    o_dn = [None, None]
    o_am = [None, None]
    cur = 0
    stride = max(1, (n + (1 << 16) - 1) >> 16)
    if W * H > (1 << 16):
        assert stride > 1
    for frame in range(3):
        print(f"[gpu] frame {frame}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        if guided:
            run.guide_refresh(guide, cams)
            o_lobes[cur] = run.lobes_host(run.cur)
            osc.bind_guide(o_lobes[cur], o_lobes[prev], o_samples)
        run.path_trace(cams, s)
        osc.gris_path_trace(cams, dn, am, s, o_resv[cur])
        assert not gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur]).any()

        run.learned_refresh(mlp, cams)
        g_learn = run.learned(run.resv[run.cur])
        offset = (cams[0].frameIndex & 0x7FFFFFFF) % stride
        new_learn, bpos, btgt = osc.learned_refresh(W, H, dn, o_resv[cur], g_learn, stride, offset)
        written = dn.reshape(n, 4)[:, 0] != 0
        o_learn[cur] = np.where(written[:, None], new_learn, o_learn[cur])
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur])
        assert not bad.any(), f"refresh: {int(bad.sum())} reservoirs differ"
        assert bits_equal(g_learn[written], o_learn[cur][written]), "learned planes differ"
        gpos, gtgt = mlp.batch()
        assert len(gpos) == len(bpos)
        assert bits_equal(gpos, bpos) and bits_equal(gtgt, btgt), "training batch differs"
        need = btgt[:, 0] >= 0
        # surface reconnection samples (the ones the network serves) are a minority where many
        # paths escape to the sky or reconnect on a light
        assert need.sum() > 0.02 * len(need), "too few training samples"
        # the predictions the GPU used at the batch pixels match the oracle network
        pix = np.arange(len(bpos)) * stride + offset
        params = mlp.params()
        want = omlp.forward(params[0], params[1], bpos[need], mode=1)
        rel = np.abs(g_learn[pix[need], :3] - want) / want
        assert np.median(rel) < 1e-4 and rel.max() < 5e-2

        run.temporal(cams, s)
        o_learn_temp_new = np.zeros_like(o_learn_temp)
        osc.gris_temporal(cams, dn, am, mv, pdn, pam, s, o_resv[cur], o_resv[prev], o_temp, o_learn[cur],
                          o_learn[prev], o_learn_temp_new)
        o_learn_temp = np.where(written[:, None], o_learn_temp_new, o_learn_temp)
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.temp), o_temp)
        assert not bad.any(), f"frame {frame} temporal: {int(bad.sum())} reservoirs differ"
        assert bits_equal(run.learned(run.temp)[written], o_learn_temp[written])

        run.spatial(cams, s)
        o_learn_new = np.zeros_like(o_learn[cur])
        osc.gris_spatial(cams, dn, am, s, o_temp, o_resv[cur], o_out, o_learn_temp, o_learn_new)
        o_learn[cur] = np.where(written[:, None], o_learn_new, o_learn[cur])
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame} spatial: {int(bad.sum())} reservoirs differ"
        assert bits_equal(run.learned(run.resv[run.cur])[written], o_learn[cur][written])
        out = run.output(1)
        d = np.abs(out.astype(np.float64) - o_out)
        ok = (d <= REL_TOL * np.maximum(np.abs(o_out), 1.0)).all(-1)
        assert ok.all(), f"frame {frame}: {int((~ok).sum())} output pixels differ"
        assert np.isfinite(out).all()

        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        rt.check(rt.core().rt_mlp_train_frame(mlp.handle, loss.data_ptr(), gpuhelp.stream()), "train frame")
        torch.cuda.synchronize()
        print(f"[gpu] frame {frame} training loss {loss.item():.4f}", flush=True)
        if guided:
            fg = dn.reshape(n, 4)[:, 0] > 0
            assert bits_equal(run.samples_host()[fg], o_samples[fg]), "training records differ"
            run.guide_train(guide, cams)
        run.flip()
        cur ^= 1
    if guided:
        osc.bind_guide()
