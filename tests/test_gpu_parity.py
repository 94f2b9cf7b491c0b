"""GPU parity: every kernel of the hot path against the oracle, through the C ABI.

Integer/index work (LBVH codes, order, topology, hit ids) must be bit-exact;
floating-point per-pixel results are compared at 1e-4 relative (the workload's
stated tolerance for the plain form) and are expected to be identical because both
sides follow one FP contract (no contraction, correctly rounded div/sqrt,
fixed sin/cos definition).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt
REL_TOL = 1e-4


def close(a, b):
    """Per-pixel agreement within REL_TOL relative (absolute below 1)."""
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    tol = REL_TOL * np.maximum(np.abs(b.astype(np.float64)), 1.0)
    return (d <= tol) | (np.isnan(a) & np.isnan(b))


@pytest.fixture(scope="module", params=[0, 2, 3, 4], ids=["boxroom", "floor", "scene225k", "scene3.0M"])
def scene(request):
    print(f"[gpu] scene config {request.param}", flush=True)
    hs = rt.HostScene(request.param, 0)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    return hs, ds, osc


def test_lbvh_bitexact(scene):
    hs, ds, osc = scene
    codes, order, nodes, info = gpuhelp.gpu_bvh(ds)
    ocodes, oorder, onodes, obox = osc.bvh()
    assert info.numPrimitives == osc.n
    assert np.array_equal(codes, ocodes), "Morton codes differ"
    assert np.array_equal(order, oorder), "radix sort order differs"
    assert np.array_equal(nodes, onodes), "LBVH nodes differ"
    assert np.array_equal(np.array(info.sceneLo, np.float32), obox[:3])
    nodes4 = gpuhelp.gpu_bvh4(ds)
    assert np.array_equal(nodes4, osc.bvh4()), "4-wide nodes differ"
    # the reported 4-wide depth is the number of levels below the root, walked on the oracle's tree
    levels, cur = 0, np.array([0])
    while cur.size:
        levels += 1
        ch = nodes4[cur, 24:28].reshape(-1)
        cur = ch[(ch != 0xFFFFFFFF) & (ch & 0x80000000 == 0)].astype(np.int64)
    assert info.bvh4Depth == levels
    assert 3 * info.bvh4Depth <= 256


def test_trace_rays_bitexact(scene):
    hs, ds, osc = scene
    lo, hi = (-1.0, 2.0) if hs.arrays.numInstances > 2 else (-3.0, 3.0)
    rays = gpuhelp.random_rays(200000, lo, hi, seed=1)
    for mode in (0, 1):
        t, bary, inst, prim = gpuhelp.gpu_trace(ds, rays, mode)
        ot, obary, oinst, oprim, _ = osc.trace(rays, mode)
        assert np.array_equal(inst, oinst), f"mode {mode}: hit instance differs"
        if mode == 0:
            assert np.array_equal(prim, oprim)
            assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(bary.view(np.uint32), obary.view(np.uint32))


def test_trace_counters_match_oracle(scene):
    hs, ds, osc = scene
    rays = gpuhelp.random_rays(20000, -1.0, 2.0, seed=2)
    rt.check(rt.core().rt_counters_enable(1), "counters")
    try:
        out = np.zeros(3, np.uint64)
        rt.check(rt.core().rt_counters_read(out.ctypes.data, 1), "read")
        gpuhelp.gpu_trace(ds, rays, 0)
        rt.check(rt.core().rt_counters_read(out.ctypes.data, 1), "read")
    finally:
        rt.core().rt_counters_enable(0)
    _, _, _, _, st = osc.trace(rays, 0)
    assert out[0] == st[1] and out[1] == st[2] and out[2] == st[0]


def test_gbuffer_bitexact(scene):
    hs, ds, osc = scene
    W, H = 96, 64
    run = gpuhelp.GrisRun(ds, W, H)
    for frame in (0, 3):
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv = run.gbuffer_host()
        odn, oam, omv, _ = osc.gbuffer(cams[0])
        assert np.array_equal(dn.view(np.uint32), odn.view(np.uint32)), "depthNormal differs"
        assert np.array_equal(am, oam), "albedoMatId differs"
        assert np.array_equal(mv, omv), "motion vectors differ"


def test_pt_plain_parity(scene):
    hs, ds, osc = scene
    W, H = 96, 64
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    run.plain(cams, 4)
    direct, indirect = run.output(0), run.output(1)
    dn, am, _ = run.gbuffer_host()
    od, oi, _ = osc.pt_plain(cams[0], dn, am, 4)
    for name, g, o in (("direct", direct, od), ("indirect", indirect, oi)):
        ok = close(g, o).all(axis=-1)
        assert ok.mean() == 1.0, f"{name}: {int((~ok).sum())} pixels outside {REL_TOL}"
        assert np.isfinite(g).all()


def run_gris_frames(hs, ds, osc, W, H, s, frames, check_passes=True):
    """Runs `frames` GRIS frames on the GPU and the oracle side by side; returns the
    number of differing reservoirs / output pixels per pass and frame."""
    run = gpuhelp.GrisRun(ds, W, H)
    n = W * H
    o_dn, o_am = [None, None], [None, None]
    o_resv = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    o_temp = orc.reset_reservoirs(n)
    o_out = np.zeros((H, W, 4), np.float32)
    cur = 0
    report = []
    for frame in range(frames):
        print(f"[gpu] frame {frame} {W}x{H}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        g_dn, g_am, g_mv = run.gbuffer_host()
        gb_bad = int((~gpuhelp.same_float(g_dn, dn)).any(-1).sum() + (g_am != am).any(-1).sum()
                     + (g_mv != mv).any(-1).sum())
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        run.path_trace(cams, s)
        osc.gris_path_trace(cams, dn, am, s, o_resv[cur])
        pt_bad = int(gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur]).sum())
        run.temporal(cams, s)
        osc.gris_temporal(cams, dn, am, mv, pdn, pam, s, o_resv[cur], o_resv[prev], o_temp)
        tp_bad = int(gpuhelp.compare_reservoirs(run.reservoirs(run.temp), o_temp).sum())
        run.spatial(cams, s)
        osc.gris_spatial(cams, dn, am, s, o_temp, o_resv[cur], o_out)
        sp_bad = int(gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur]).sum())
        out = run.output(1)
        out_bad = int((~close(out, o_out).all(axis=-1)).sum())
        report.append((gb_bad, pt_bad, tp_bad, sp_bad, out_bad))
        run.flip()
        cur ^= 1
    return report


def test_gris_terrain_1080p_parity():
    """The benchmark workload itself: 1920x1080, 1.11M triangles, hybrid shift, temporal +
    spatial (3 neighbours, 30 px, cap 20), depth 6 — two frames, every pass against the oracle."""
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    s = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=6, neighbors=3, radius=30.0)
    report = run_gris_frames(hs, ds, osc, 1920, 1080, s, 2)
    print("[gpu] mismatches per frame (gbuffer, path trace, temporal, spatial, output):", report, flush=True)
    assert all(sum(r) == 0 for r in report), report


def test_pt_plain_boxroom_256():
    """Plain PT with NEE+MIS on the box room at 256x256, depth 4 (the plain-PT workload)."""
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    W = H = 256
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    run.plain(cams, 4)
    dn, am, _ = run.gbuffer_host()
    od, oi, _ = osc.pt_plain(cams[0], dn, am, 4)
    for g, o in ((run.output(0), od), (run.output(1), oi)):
        assert close(g, o).all()


def test_gpu_frames_are_deterministic():
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 640, 360
    s = rt.settings(temporal=True, spatial=True, max_depth=6, radius=10.0)
    outs = []
    for _ in range(2):
        run = gpuhelp.GrisRun(ds, W, H)
        for frame in range(3):
            out = run.frame(hs.camera(frame, W, H), s)
            run.flip()
        outs.append((out, run.reservoirs(run.resv[run.cur ^ 1])))
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert not gpuhelp.compare_reservoirs(outs[0][1], outs[1][1]).any()
    assert np.isfinite(outs[0][0]).all() and (outs[0][0][..., :3] >= 0).all()


@pytest.mark.parametrize("shift,unbiased", [(2, False), (0, False), (2, True), (0, True)],
                         ids=["hybrid", "reconnection", "hybrid-unbiased", "reconnection-unbiased"])
def test_gris_sequence_parity(scene, shift, unbiased):
    """Three frames of the full GRIS pipeline (temporal + spatial), pass by pass; also in the
    unbiased reuse mode (1/Z weights, consistent shifted contributions, light MIS re-evaluated)."""
    hs, ds, osc = scene
    W, H = 80, 48
    s = rt.settings(shift=shift, temporal=True, spatial=True, max_depth=6, cap=20, radius=20.0, unbiased=unbiased)
    run = gpuhelp.GrisRun(ds, W, H)
    n = W * H
    o_dn = [None, None]
    o_am = [None, None]
    o_resv = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    o_temp = orc.reset_reservoirs(n)
    o_out = np.zeros((H, W, 4), np.float32)
    cur = 0
    for frame in range(3):
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        if o_dn[prev] is None:  # first frame: the GPU's previous parity is zero-filled too
            pass
        run.path_trace(cams, s)
        osc.gris_path_trace(cams, dn, am, s, o_resv[cur])
        bad = gpuhelp.compare_reservoirs(run.reservoirs(run.resv[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame} path trace: {int(bad.sum())} reservoirs differ"
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
        assert ok.all(), f"frame {frame}: {int((~ok).sum())} output pixels outside {REL_TOL}"
        run.flip()
        cur ^= 1


def test_nonfinite_rays_miss_without_traversal():
    """Rays with NaN or infinite components miss at once on the GPU as in the oracle (before the
    fix, such a ray walked the whole 1.1M-triangle tree: ~1 s for one lane)."""
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    rays = gpuhelp.random_rays(4096, -8.0, 8.0, seed=4)
    rays[:, 2] += 6.0
    bad = np.arange(0, 4096, 3)
    rays[bad, 3 + (bad % 3)] = np.where(bad % 2 == 0, np.nan, np.inf)
    rt.check(rt.core().rt_counters_enable(1), "counters")
    try:
        out = np.zeros(3, np.uint64)
        rt.check(rt.core().rt_counters_read(out.ctypes.data, 1), "read")
        t, bary, inst, prim = gpuhelp.gpu_trace(ds, rays[bad], 0)
        rt.check(rt.core().rt_counters_read(out.ctypes.data, 1), "read")
    finally:
        rt.core().rt_counters_enable(0)
    assert (inst == 0xFFFFFFFF).all() and out[0] == 0 and out[1] == 0 and out[2] == len(bad)
    for mode in (0, 1):
        t, bary, inst, prim = gpuhelp.gpu_trace(ds, rays, mode)
        ot, obary, oinst, oprim, _ = osc.trace(rays, mode)
        assert np.array_equal(inst, oinst)
        if mode == 0:
            assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))


@pytest.mark.parametrize("method", ["plain", "gris"])
def test_converged_accumulation_matches_oracle(method):
    """Progressive accumulation over 4096 frames (4096 spp) of a still view on the GPU (the
    renderer's FP32 (acc * k + c) / (k + 1) with k = frameIndex) against the oracle's
    FP64-accumulated converged image (the mean of its 4096 per-frame estimates, summed in float64;
    the same seeds, so every frame's estimate is the GPU's). Stated bound: relMSE < 1e-9, the FP32
    accumulation's rounding."""
    hs = rt.HostScene(0, 4)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    W, H = 48, 32
    base = hs.camera(0, W, H)
    run = gpuhelp.GrisRun(ds, W, H)
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=4, cap=20, radius=10.0)
    n = W * H
    o_direct = np.zeros((H, W, 4), np.float32)
    o_indirect = np.zeros((H, W, 4), np.float32)
    acc_direct = np.zeros((H, W, 3), np.float64)
    acc_indirect = np.zeros((H, W, 3), np.float64)
    o_resv = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    o_temp = orc.reset_reservoirs(n)
    o_dn = [None, None]
    o_am = [None, None]
    cur = 0
    frames = 4096
    for i in range(frames):
        if i % 512 == 0:
            print(f"[gpu] converged {method}: frame {i}", flush=True)
        cams = gpuhelp.still_cams(base, i, 5)
        ocams = gpuhelp.still_cams(base, i, 5)
        if i > 0:   # the oracle writes each frame's own estimate (k = 0); the clear flag stays on frame 0
            ocams[0].frameIndex = ocams[1].frameIndex = 0
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(ocams[0])
        if method == "plain":
            run.plain(cams, 4)
            osc.pt_plain(ocams[0], dn, am, 4, o_direct, o_indirect)
            acc_direct += o_direct[..., :3]
        else:
            o_dn[cur], o_am[cur] = dn, am
            prev = cur ^ 1
            pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
            pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
            run.path_trace(cams, s)
            run.temporal(cams, s)
            run.spatial(cams, s)
            osc.gris_path_trace(ocams, dn, am, s, o_resv[cur])
            osc.gris_temporal(ocams, dn, am, mv, pdn, pam, s, o_resv[cur], o_resv[prev], o_temp)
            osc.gris_spatial(ocams, dn, am, s, o_temp, o_resv[cur], o_indirect)
            run.flip()
            cur ^= 1
        acc_indirect += o_indirect[..., :3]
    outs = [(run.output(1), acc_indirect / frames)] + ([(run.output(0), acc_direct / frames)] if method == "plain" else [])
    for g, o in outs:
        rel = (((g[..., :3].astype(np.float64) - o) ** 2) / (o ** 2 + 1e-2)).sum(-1).mean()
        print(f"[gpu] {method}: converged relMSE GPU (FP32 accumulation) vs FP64 oracle {rel:.3e}", flush=True)
        assert rel < 1e-9
        assert o.mean() > 0
