"""Batched views against per-view launches: views of two scenes with different sizes
and depths rendered together (one G-buffer, one persistent path-trace, one temporal and
one spatial launch per frame) must give the same G-buffers, reservoirs and outputs, bit
for bit, as each view rendered on its own; and the autotuner returns a valid tuning."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402

rt = gpuhelp.rt

# (scene config, width, height, max depth)
VIEWS = [(0, 96, 64, 4), (2, 40, 24, 3), (0, 130, 71, 8), (2, 64, 64, 6), (0, 8, 8, 5)]


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
        for frame in range(3):
            print(f"[gpu] batch frame {frame}", flush=True)
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
            views[i] = make_view(batched[i], ds[c], hs[c].camera(3, W, H), sets[i])
        best = rt.BatchTuning()
        ms = C.c_float()
        rt.check(L.rt_batch_autotune(b, views, len(VIEWS), 2, gpuhelp.stream(), C.byref(best), C.byref(ms)), "tune")
        print(f"[gpu] autotune: groups {best.persistentGroups} swizzle {best.xcdSwizzle} {ms.value:.3f} ms", flush=True)
        assert best.persistentGroups > 0 and best.xcdSwizzle in (0, 1) and 0 < ms.value < 1e4
    finally:
        L.rt_batch_destroy(b)
