"""Batched views against per-view launches: views of two scenes with different sizes
and depths rendered together (one G-buffer, one persistent path-trace, one temporal and
one spatial launch per frame) must give the same G-buffers, reservoirs and outputs, bit
for bit, as each view rendered on its own, under every tuning (register budget, queue
tile shape, persistent grid, eager launches or graph replay); and the autotuner returns
a valid tuning."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402

rt = gpuhelp.rt

# (scene config, width, height, max depth)
VIEWS = [(0, 96, 64, 4), (2, 40, 24, 3), (0, 130, 71, 8), (2, 64, 64, 6), (0, 8, 8, 5)]

# per frame: (persistentGroups, xcdSwizzle, waves, queueTile, graph, shading); the third
# and fourth frames replay the graph built by the second with updated kernel arguments;
# the last four run the wavefront (material-sorted) variant, with 7 workgroups (many
# trips of its grid-stride loops) and as a graph built and then replayed
TUNINGS = [(0, 16, 0, 0, 0, 0), (0, 16, 3, 1, 1, 0), (7, 16, 3, 1, 1, 0), (0, 0, 5, 2, 1, 0), (0, 128, 4, 0, 0, 1),
           (7, 16, 4, 0, 0, 1), (700, 16, 4, 0, 1, 1), (700, 16, 4, 0, 1, 1)]


def tuning(persistent, swizzle, waves, tile, graph, shading=0):
    t = rt.BatchTuning()
    t.persistentGroups, t.xcdSwizzle, t.waves, t.queueTile, t.graph = persistent, swizzle, waves, tile, graph
    t.shading = shading
    return t


def make_view(run, sc, cams, s):
    v = rt.View()
    v.scene = sc.handle.value if hasattr(sc.handle, "value") else sc.handle
    C.memmove(C.addressof(v.cams), C.addressof(cams), C.sizeof(rt.Camera) * 2)
    v.frame = run.frame_struct()
    v.settings = s
    return v


def test_batched_views_match_single_views():
    hs = {0: rt.HostScene(0, 3), 2: rt.HostScene(2, 3)}
    ds = {k: rt.DeviceScene(h.arrays) for k, h in hs.items()}
    single = [gpuhelp.GrisRun(ds[c], W, H) for c, W, H, _ in VIEWS]
    batched = [gpuhelp.GrisRun(ds[c], W, H) for c, W, H, _ in VIEWS]
    sets = [rt.settings(shift=2, temporal=True, spatial=True, max_depth=d, cap=20, radius=10.0) for _, _, _, d in VIEWS]
    b = C.c_void_p()
    L = rt.core()
    rt.check(L.rt_batch_create(len(VIEWS), C.byref(b)), "batch create")
    try:
        for frame in range(len(TUNINGS)):
            print(f"[gpu] batch frame {frame} tuning {TUNINGS[frame]}", flush=True)
            rt.check(L.rt_batch_set_tuning(b, C.byref(tuning(*TUNINGS[frame]))), "tuning")
            views = (rt.View * len(VIEWS))()
            for i, (c, W, H, _) in enumerate(VIEWS):
                cams = hs[c].camera(frame, W, H)
                r = single[i]
                r.gbuffer(cams[0])
                r.path_trace(cams, sets[i])
                r.temporal(cams, sets[i])
                r.spatial(cams, sets[i])
                views[i] = make_view(batched[i], ds[c], cams, sets[i])
            rt.check(L.rt_batch_frame(b, views, len(VIEWS), gpuhelp.stream()), "batch frame")
            torch.cuda.synchronize()
            for i, (c, W, H, _) in enumerate(VIEWS):
                a, z = single[i], batched[i]
                for ga, gz in zip(a.gbuffer_host(), z.gbuffer_host()):
                    assert np.array_equal(ga.view(np.uint8), gz.view(np.uint8)), f"view {i}: G-buffer differs"
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.resv[a.cur]), z.reservoirs(z.resv[z.cur])).any()
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.temp), z.reservoirs(z.temp)).any()
                assert np.array_equal(a.output(1).view(np.uint32), z.output(1).view(np.uint32)), f"view {i}: output"
                a.flip()
                z.flip()
        # autotuner: a valid tuning and a finite time
        views = (rt.View * len(VIEWS))()
        for i, (c, W, H, _) in enumerate(VIEWS):
            views[i] = make_view(batched[i], ds[c], hs[c].camera(len(TUNINGS), W, H), sets[i])
        best = rt.BatchTuning()
        ms = C.c_float()
        rt.check(L.rt_batch_autotune(b, views, len(VIEWS), 2, gpuhelp.stream(), C.byref(best), C.byref(ms)), "tune")
        print(f"[gpu] autotune: groups {best.persistentGroups} swizzle {best.xcdSwizzle} waves {best.waves} "
              f"tile {best.queueTile} graph {best.graph} shading {best.shading} {ms.value:.3f} ms", flush=True)
        assert best.xcdSwizzle in (0, 16, 128) and best.waves in (3, 4, 5) and best.queueTile in (0, 1, 2)
        assert best.graph in (0, 1) and best.shading in (0, 1) and 0 < ms.value < 1e4
        got = rt.BatchTuning()
        rt.check(L.rt_batch_get_tuning(b, C.byref(got)), "get tuning")
        assert bytes(got) == bytes(best)
        # invalid tunings are rejected
        for bad in ((0, 16, 6, 0, 0), (0, 16, 4, 3, 0), (0, 16, 4, 0, 2), (0, 16, 4, 0, 0, 2)):
            assert L.rt_batch_set_tuning(b, C.byref(tuning(*bad))) != 0
    finally:
        L.rt_batch_destroy(b)


def _handle(h):
    return h.value if hasattr(h, "value") else h


# (scene config, width, height, max depth, learned target, unbiased reuse, learned proposal)
FORM_VIEWS = [(1, 96, 64, 6, True, True, True), (0, 40, 24, 3, False, False, False), (1, 64, 48, 5, True, False, False),
              (0, 48, 40, 4, False, True, True), (1, 33, 17, 6, False, False, True)]


def test_batched_learned_forms_match_single_views():
    """Views of every form in one batch (learned target, unbiased reuse and learned proposal in
    several combinations, next to a plain view), each with its own networks: G-buffers,
    reservoirs with their learned planes, lobe planes, training records, outputs and the
    networks' parameters after three frames equal those of the same views rendered alone."""
    from test_gpu_guide import GuidedRun
    from test_gpu_learned import scene_bounds

    hs = {c: rt.HostScene(c, 1) for c in (0, 1)}
    ds = {c: rt.DeviceScene(h.arrays) for c, h in hs.items()}
    bounds = {c: scene_bounds(d) for c, d in ds.items()}

    def nets(i, c, learned, guided):
        lo, hi = bounds[c]
        t = rt.Mlp(rt.mlp_config(lo, hi, log2_table=15, seed=10 + i)) if learned else None
        g = rt.Mlp(rt.mlp_config(lo, hi, log2_table=15, seed=40 + i), guide=True) if guided else None
        return t, g

    single = [GuidedRun(ds[v[0]], v[1], v[2]) for v in FORM_VIEWS]
    batched = [GuidedRun(ds[v[0]], v[1], v[2]) for v in FORM_VIEWS]
    s_nets = [nets(i, v[0], v[4], v[6]) for i, v in enumerate(FORM_VIEWS)]
    b_nets = [nets(i, v[0], v[4], v[6]) for i, v in enumerate(FORM_VIEWS)]
    sets = [rt.settings(shift=2, temporal=True, spatial=True, max_depth=d, cap=20, radius=10.0, learned=lt,
                        unbiased=ub, guided=gd) for _, _, _, d, lt, ub, gd in FORM_VIEWS]
    b = C.c_void_p()
    L = rt.core()
    rt.check(L.rt_batch_create(len(FORM_VIEWS), C.byref(b)), "batch create")
    assert L.rt_batch_set_tuning(b, C.byref(tuning(0, 5000, 0, 0, 0))) != 0  # more than 4096 tiles per XCD chunk
    try:
        for frame in range(3):
            # runs of 3 tiles per XCD (partial last runs); graph replay from the second frame on
            rt.check(L.rt_batch_set_tuning(b, C.byref(tuning(0, 3, 4, frame, int(frame > 0)))), "tuning")
            print(f"[gpu] batched forms frame {frame}", flush=True)
            views = (rt.View * len(FORM_VIEWS))()
            for i, (c, W, H, _, lt, ub, gd) in enumerate(FORM_VIEWS):
                cams = hs[c].camera(frame, W, H)
                r, (t, g) = single[i], s_nets[i]
                r.gbuffer(cams[0])
                if gd:
                    r.guide_refresh(g, cams)
                r.path_trace(cams, sets[i])
                if lt:
                    r.learned_refresh(t, cams)
                r.temporal(cams, sets[i])
                r.spatial(cams, sets[i])
                if lt:
                    rt.check(L.rt_mlp_train_frame(t.handle, None, gpuhelp.stream()), "train frame")
                if gd:
                    r.guide_train(g, cams)
                v = make_view(batched[i], ds[c], cams, sets[i])
                bt, bg = b_nets[i]
                v.target = _handle(bt.handle) if bt else None
                v.guide = _handle(bg.handle) if bg else None
                views[i] = v
            rt.check(L.rt_batch_frame(b, views, len(FORM_VIEWS), gpuhelp.stream()), "batch frame")
            torch.cuda.synchronize()
            for i, (c, W, H, _, lt, ub, gd) in enumerate(FORM_VIEWS):
                a, z = single[i], batched[i]
                for ga, gz in zip(a.gbuffer_host(), z.gbuffer_host()):
                    assert np.array_equal(ga.view(np.uint8), gz.view(np.uint8)), f"view {i}: G-buffer differs"
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.resv[a.cur]), z.reservoirs(z.resv[z.cur])).any()
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.temp), z.reservoirs(z.temp)).any()
                if lt:
                    assert np.array_equal(a.learned(a.resv[a.cur]).view(np.uint32),
                                          z.learned(z.resv[z.cur]).view(np.uint32)), f"view {i}: learned planes"
                if gd:
                    assert np.array_equal(a.lobes_host(a.cur).view(np.uint32),
                                          z.lobes_host(z.cur).view(np.uint32)), f"view {i}: lobe planes"
                    assert np.array_equal(a.samples_host().view(np.uint32),
                                          z.samples_host().view(np.uint32)), f"view {i}: training records"
                assert np.array_equal(a.output(1).view(np.uint32), z.output(1).view(np.uint32)), f"view {i}: output"
                for na, nz in zip(s_nets[i], b_nets[i]):
                    if na is not None:
                        for pa, pz in zip(na.params(), nz.params()):
                            assert np.array_equal(pa.view(np.uint32), pz.view(np.uint32)), f"view {i}: network"
                a.flip()
                z.flip()
        # a network bound to two views of one frame is rejected before anything runs
        views = (rt.View * 2)()
        for k in range(2):
            c, W, H = FORM_VIEWS[0][:3]
            v = make_view(batched[0], ds[c], hs[c].camera(3, W, H), sets[0])
            v.target = _handle(b_nets[0][0].handle)
            v.guide = _handle(b_nets[0][1].handle)
            views[k] = v
        assert L.rt_batch_frame(b, views, 2, gpuhelp.stream()) != 0
        assert b"more than one view" in rt.core().rt_last_error()
        # the wavefront variant does not run the learned proposal: refused, not silently replaced
        rt.check(L.rt_batch_set_tuning(b, C.byref(tuning(0, 16, 4, 0, 0, 1))), "tuning")
        views = (rt.View * 1)()
        c, W, H = FORM_VIEWS[0][:3]
        v = make_view(batched[0], ds[c], hs[c].camera(3, W, H), sets[0])
        v.target, v.guide = _handle(b_nets[0][0].handle), _handle(b_nets[0][1].handle)
        views[0] = v
        assert L.rt_batch_frame(b, views, 1, gpuhelp.stream()) != 0
        assert b"learned proposal" in rt.core().rt_last_error()
    finally:
        L.rt_batch_destroy(b)


# (scene config, width, height, max depth, learned target, unbiased reuse)
WF_VIEWS = [(1, 72, 40, 6, True, False), (0, 40, 24, 5, False, True), (1, 33, 17, 7, True, True), (0, 24, 24, 3, False, False)]


def test_wavefront_learned_and_unbiased_views():
    """The wavefront variant beside views with the learned target and the unbiased weights, eager
    and as a graph: every view equals the same view rendered alone, bit for bit."""
    from test_gpu_learned import scene_bounds

    hs = {c: rt.HostScene(c, 1) for c in (0, 1)}
    ds = {c: rt.DeviceScene(h.arrays) for c, h in hs.items()}
    single = [gpuhelp.GrisRun(ds[v[0]], v[1], v[2]) for v in WF_VIEWS]
    batched = [gpuhelp.GrisRun(ds[v[0]], v[1], v[2]) for v in WF_VIEWS]

    def net(i, c, learned):
        lo, hi = scene_bounds(ds[c])
        return rt.Mlp(rt.mlp_config(lo, hi, log2_table=15, seed=70 + i)) if learned else None

    s_nets = [net(i, v[0], v[4]) for i, v in enumerate(WF_VIEWS)]
    b_nets = [net(i, v[0], v[4]) for i, v in enumerate(WF_VIEWS)]
    sets = [rt.settings(shift=2, temporal=True, spatial=True, max_depth=d, cap=20, radius=10.0, learned=lt, unbiased=ub)
            for _, _, _, d, lt, ub in WF_VIEWS]
    b = C.c_void_p()
    L = rt.core()
    rt.check(L.rt_batch_create(len(WF_VIEWS), C.byref(b)), "batch create")
    try:
        for frame in range(3):
            rt.check(L.rt_batch_set_tuning(b, C.byref(tuning(64, 16, 4, 0, int(frame > 0), 1))), "tuning")
            print(f"[gpu] wavefront forms frame {frame}", flush=True)
            views = (rt.View * len(WF_VIEWS))()
            for i, (c, W, H, _, lt, ub) in enumerate(WF_VIEWS):
                cams = hs[c].camera(frame, W, H)
                r, t = single[i], s_nets[i]
                r.gbuffer(cams[0])
                r.path_trace(cams, sets[i])
                if lt:
                    r.learned_refresh(t, cams)
                r.temporal(cams, sets[i])
                r.spatial(cams, sets[i])
                if lt:
                    rt.check(L.rt_mlp_train_frame(t.handle, None, gpuhelp.stream()), "train frame")
                v = make_view(batched[i], ds[c], cams, sets[i])
                v.target = _handle(b_nets[i].handle) if lt else None
                views[i] = v
            rt.check(L.rt_batch_frame(b, views, len(WF_VIEWS), gpuhelp.stream()), "batch frame")
            torch.cuda.synchronize()
            for i, (c, W, H, _, lt, ub) in enumerate(WF_VIEWS):
                a, z = single[i], batched[i]
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.resv[a.cur]), z.reservoirs(z.resv[z.cur])).any()
                assert not gpuhelp.compare_reservoirs(a.reservoirs(a.temp), z.reservoirs(z.temp)).any()
                if lt:
                    assert np.array_equal(a.learned(a.resv[a.cur]).view(np.uint32),
                                          z.learned(z.resv[z.cur]).view(np.uint32)), f"view {i}: learned planes"
                assert np.array_equal(a.output(1).view(np.uint32), z.output(1).view(np.uint32)), f"view {i}: output"
                a.flip()
                z.flip()
    finally:
        L.rt_batch_destroy(b)


def same_bytes(a, b):
    """Byte equality of two device buffers, on the device."""
    return bool(torch.equal(a, b))


def assert_views_equal(i, a, z, guided=False):
    c = a.cur
    for name, x, y in (("depthNormal", a.dn[c], z.dn[c]), ("albedoMatId", a.am[c], z.am[c]), ("motion", a.mv, z.mv),
                       ("reservoirs", a.resv[c], z.resv[c]), ("temp reservoirs", a.temp, z.temp),
                       ("indirect output", a.indirect, z.indirect)):
        assert same_bytes(x, y), f"view {i}: {name} differ"
    if guided:
        assert same_bytes(a.lobes[c], z.lobes[c]), f"view {i}: lobe planes differ"
        assert same_bytes(a.samples, z.samples), f"view {i}: training records differ"


def config2_specs():
    """The batched workload's views (as bench.py draws them): 64 square views, sizes log-uniform in
    [128, 1024] (multiples of 8), depth U{3..8}, scenes round robin, a camera-walk offset each."""
    g = np.random.default_rng(2)
    specs = []
    for i in range(64):
        size = int(np.exp(g.uniform(np.log(128), np.log(1024)))) // 8 * 8
        specs.append((i % 4, size, int(g.integers(3, 9)), int(g.integers(0, 64))))
    return specs


def test_config2_batch_matches_single_views():
    """The batched workload at full size: 64 views (16.0M pixels) over the four scenes of ~20k, ~225k,
    ~1.1M and ~3.0M triangles in one rt_batch_frame per frame, against the same views rendered one by
    one: G-buffers, reservoirs and outputs byte-identical over two frames."""
    hosts = [rt.HostScene(0, 2), rt.HostScene(3, 2), rt.HostScene(1, 1), rt.HostScene(4, 2)]
    devs = [rt.DeviceScene(h.arrays) for h in hosts]
    specs = config2_specs()
    assert sum(n * n for _, n, _, _ in specs) > 15_000_000
    single = [gpuhelp.GrisRun(devs[si], n, n) for si, n, _, _ in specs]
    batched = [gpuhelp.GrisRun(devs[si], n, n) for si, n, _, _ in specs]
    sets = [rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=d, neighbors=3, radius=20.0)
            for _, _, d, _ in specs]
    L = rt.core()
    b = C.c_void_p()
    rt.check(L.rt_batch_create(len(specs), C.byref(b)), "batch create")
    try:
        for frame in range(2):
            print(f"[gpu] config-2 batch frame {frame}", flush=True)
            views = (rt.View * len(specs))()
            for i, (si, n, _, f0) in enumerate(specs):
                cams = hosts[si].camera((f0 + frame) % 64, n, n)
                r = single[i]
                r.gbuffer(cams[0])
                r.path_trace(cams, sets[i])
                r.temporal(cams, sets[i])
                r.spatial(cams, sets[i])
                views[i] = make_view(batched[i], devs[si], cams, sets[i])
            rt.check(L.rt_batch_frame(b, views, len(specs), gpuhelp.stream()), "batch frame")
            torch.cuda.synchronize()
            for i in range(len(specs)):
                assert_views_equal(i, single[i], batched[i])
                single[i].flip()
                batched[i].flip()
    finally:
        L.rt_batch_destroy(b)


def test_full_form_4k_batch_matches_single_views():
    """The full-form workload at full size: a 3840x2160 view of the ~4.8M-triangle scene at depth 8
    with the learned target, the learned proposal and unbiased reuse (each network trained every
    frame) and 16 plain 512x512 views in one rt_batch_frame, against the same views rendered alone:
    G-buffers, reservoirs, lobe planes, training records, outputs and both networks' parameters
    byte-identical over two frames; the 4K image is finite and non-negative and its reservoirs
    respect the M-cap."""
    from test_gpu_guide import GuidedRun
    from test_gpu_learned import scene_bounds

    hs = rt.HostScene(5, 4)
    ds = rt.DeviceScene(hs.arrays)
    assert ds.info().numPrimitives > 4_500_000
    W, H = 3840, 2160
    lo, hi = scene_bounds(ds)
    nets = [(rt.Mlp(rt.mlp_config(lo, hi, log2_table=19, seed=4)),
             rt.Mlp(rt.mlp_config(lo, hi, log2_table=19, seed=5), guide=True)) for _ in range(2)]
    s4k = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=8, neighbors=3, radius=30.0,
                      learned=True, unbiased=True, guided=True)
    s512 = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=8, neighbors=3, radius=30.0)
    offsets = np.random.default_rng(4).integers(0, 64, 16)
    single = [GuidedRun(ds, W, H)] + [gpuhelp.GrisRun(ds, 512, 512) for _ in range(16)]
    batched = [GuidedRun(ds, W, H)] + [gpuhelp.GrisRun(ds, 512, 512) for _ in range(16)]
    L = rt.core()
    b = C.c_void_p()
    rt.check(L.rt_batch_create(17, C.byref(b)), "batch create")
    try:
        for frame in range(2):
            print(f"[gpu] 4K full-form batch frame {frame}", flush=True)
            views = (rt.View * 17)()
            for i in range(17):
                r = single[i]
                if i == 0:
                    cams = hs.camera(frame, W, H)
                    mlp, guide = nets[0]
                    r.gbuffer(cams[0])
                    r.guide_refresh(guide, cams)
                    r.path_trace(cams, s4k)
                    r.learned_refresh(mlp, cams)
                    r.temporal(cams, s4k)
                    r.spatial(cams, s4k)
                    rt.check(L.rt_mlp_train_frame(mlp.handle, None, gpuhelp.stream()), "train frame")
                    r.guide_train(guide, cams)
                    v = make_view(batched[0], ds, cams, s4k)
                    v.target, v.guide = _handle(nets[1][0].handle), _handle(nets[1][1].handle)
                else:
                    cams = hs.camera(int(offsets[i - 1] + frame) % 64, 512, 512)
                    r.frame(cams, s512)
                    v = make_view(batched[i], ds, cams, s512)
                views[i] = v
            rt.check(L.rt_batch_frame(b, views, 17, gpuhelp.stream()), "batch frame")
            torch.cuda.synchronize()
            for i in range(17):
                assert_views_equal(i, single[i], batched[i], guided=(i == 0))
            for na, nz in zip(nets[0], nets[1]):
                for pa, pz in zip(na.params(), nz.params()):
                    assert np.array_equal(pa.view(np.uint32), pz.view(np.uint32)), "network parameters differ"
            out = batched[0].output(1)
            assert np.isfinite(out).all() and (out[..., :3] >= 0).all()
            assert out[..., :3].mean() > 0
            m = batched[0].reservoirs(batched[0].resv[batched[0].cur])["sampleCount"]
            assert m.max() <= 20
            for r in single + batched:
                r.flip()
    finally:
        L.rt_batch_destroy(b)
