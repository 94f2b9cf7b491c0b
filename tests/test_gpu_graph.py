"""Frame graphs against eager launches: the same frames run eagerly on one set of
buffers and as graph replays on another must give bit-identical reservoirs, learned
planes, outputs and network parameters, frame after frame, with the program's own
eager GPU work and synchronisations between the replays."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402

rt = gpuhelp.rt


def frames_of(run):
    """rt_frame bindings of both parities of a GrisRun."""
    fs = (rt.Frame * 2)()
    cur = run.cur
    for j in range(2):
        run.cur = j
        fs[j] = run.frame_struct()
    run.cur = cur
    return fs


def scene_bounds(ds):
    i = ds.info()
    lo, hi = np.array(i.sceneLo), np.array(i.sceneHi)
    ext = np.maximum(hi - lo, 1e-3)
    return tuple(lo - 0.1 * ext), tuple(hi + 0.1 * ext)


@pytest.mark.parametrize("learned,unbiased,guided", [(False, False, False), (True, False, False), (True, True, True)],
                         ids=["luminance", "learned", "learned-unbiased-proposal"])
def test_frame_graph_matches_eager(learned, unbiased, guided):
    """With guided: the proposal's refresh, gather and training step are graph nodes too."""
    from test_gpu_guide import GuidedRun

    print(f"[gpu] frame graph learned={learned} unbiased={unbiased} proposal={guided}", flush=True)
    hs = rt.HostScene(0, 2)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 96, 64
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=6, cap=20, radius=20.0, learned=learned,
                    unbiased=unbiased, guided=guided)
    eager, graphed = GuidedRun(ds, W, H), GuidedRun(ds, W, H)
    lo, hi = scene_bounds(ds)
    mlps = [rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=4)) for _ in range(2)] if learned else [None, None]
    guides = ([rt.Mlp(rt.mlp_config(lo, hi, log2_table=16, seed=6), guide=True) for _ in range(2)] if guided
              else [None, None])
    g = C.c_void_p()
    rt.check(rt.core().rt_frame_graph_create(ds.handle, frames_of(graphed), C.byref(s),
                                             mlps[1].handle if learned else None,
                                             guides[1].handle if guided else None, C.byref(g)), "graph create")
    L = rt.core()
    try:
        for frame in range(6):
            print(f"[gpu] frame {frame}", flush=True)
            cams = hs.camera(frame, W, H)
            # eager
            eager.gbuffer(cams[0])
            if guided:
                eager.guide_refresh(guides[0], cams)
            eager.path_trace(cams, s)
            if learned:
                eager.learned_refresh(mlps[0], cams)
            eager.temporal(cams, s)
            eager.spatial(cams, s)
            if learned:
                rt.check(L.rt_mlp_train_frame(mlps[0].handle, None, gpuhelp.stream()), "train")
            if guided:
                eager.guide_train(guides[0], cams)
            # the program's own work between replays
            junk = torch.randn(1 << 20, device="cuda").sum()
            torch.cuda.synchronize()
            assert torch.isfinite(junk)
            # graph replay
            rt.check(L.rt_frame_graph_launch(g, graphed.cur, cams, gpuhelp.stream()), "graph launch")
            torch.cuda.synchronize()
            for name, a, b in (("reservoirs", eager.resv[eager.cur], graphed.resv[graphed.cur]),
                               ("temp", eager.temp, graphed.temp)):
                ra, rb = eager.reservoirs(a), graphed.reservoirs(b)
                assert not gpuhelp.compare_reservoirs(ra, rb).any(), f"frame {frame}: {name} differ"
                if learned:
                    assert np.array_equal(eager.learned(a).view(np.uint32), graphed.learned(b).view(np.uint32))
            oa, ob = eager.output(1), graphed.output(1)
            assert np.array_equal(oa.view(np.uint32), ob.view(np.uint32)), f"frame {frame}: outputs differ"
            for nets in (mlps, guides):
                if nets[0] is not None:
                    pa, pb = nets[0].params(), nets[1].params()
                    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1]), "network parameters differ"
            if guided:
                assert np.array_equal(eager.lobes_host(eager.cur).view(np.uint32),
                                      graphed.lobes_host(graphed.cur).view(np.uint32)), f"frame {frame}: lobes"
                assert np.array_equal(eager.samples_host().view(np.uint32),
                                      graphed.samples_host().view(np.uint32)), f"frame {frame}: records"
            eager.flip()
            graphed.flip()
    finally:
        L.rt_frame_graph_destroy(g)


def test_training_is_deterministic():
    """Two networks trained on the same batches stay bit-identical (fixed-point grid gradients,
    fixed-order weight-gradient reduction)."""
    lo, hi = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    a = rt.Mlp(rt.mlp_config(lo, hi, log2_table=14, seed=2))
    b = rt.Mlp(rt.mlp_config(lo, hi, log2_table=14, seed=2))
    g = np.random.default_rng(0)
    for _ in range(20):
        pos = torch.from_numpy(g.uniform(-1, 1, (20000, 3)).astype(np.float32)).cuda()
        tgt = torch.from_numpy(np.exp(g.normal(0, 1, (20000, 3))).astype(np.float32)).cuda()
        for m in (a, b):
            rt.check(rt.core().rt_mlp_train_step(m.handle, pos.data_ptr(), tgt.data_ptr(), 20000, None,
                                                 gpuhelp.stream()), "train")
    torch.cuda.synchronize()
    pa, pb = a.params(), b.params()
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
