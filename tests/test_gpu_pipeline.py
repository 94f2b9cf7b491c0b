"""rt_pipeline (GRIS frames overlapped on two streams) against rt_renderer (one frame after the
other): over a camera walk both must leave the same final reservoirs for the last two frames and
the same indirect output, bit for bit, with biased and unbiased resampling weights, with one stream
for both, with a capped path-trace launch, and with the naive direct light or ReSTIR DI (the same
direct output and DI reservoirs); a reset restarts the walk; settings the pipeline does not take
are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt
W, H = 160, 96
FRAMES = 7
INVALID_ARGUMENT = 1  # RT_ERR_INVALID_ARGUMENT


def download_di(ptr):
    out = np.zeros(W * H, orc.DI_DTYPE)
    torch.cuda.synchronize()
    rt.check(rt.core().rt_di_reservoir_download(ptr, W, H, out.ctypes.data, gpuhelp.stream()), "di download")
    return out


def download(ptr):
    out = np.zeros(W * H, orc.RESV_DTYPE)
    torch.cuda.synchronize()
    rt.check(rt.core().rt_gris_reservoir_download(ptr, W, H, out.ctypes.data, gpuhelp.stream()), "download")
    return out


def sequential(hs, ds, s, frames, direct=0, di=None):
    """rt_renderer frames (direct method `direct`, ReSTIR PT indirect): the last frame's and the
    frame before's final reservoirs, the indirect and the direct output, and with ReSTIR DI its
    last frame's reservoirs."""
    Hh = rt.host()
    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
    try:
        rt.check(Hh.rt_renderer_set_settings(r, C.byref(s)), "settings")
        if di is not None:
            rt.check(Hh.rt_renderer_set_di_settings(r, C.byref(di)), "di settings")
        for k in range(frames):
            rt.check(Hh.rt_renderer_frame_methods(r, hs.camera(k, W, H), direct, 3, gpuhelp.stream()), "frame")
        f = rt.Frame()
        rt.check(Hh.rt_renderer_bindings(r, C.byref(f)), "bindings")
        out = np.zeros((H, W, 4), np.float32)
        rt.check(Hh.rt_renderer_read_output(r, 1, out.ctypes.data, gpuhelp.stream()), "read")
        dout = np.zeros((H, W, 4), np.float32)
        rt.check(Hh.rt_renderer_read_output(r, 0, dout.ctypes.data, gpuhelp.stream()), "read")
        dres = None
        if direct == 2:
            d = rt.DIReservoirs()
            rt.check(Hh.rt_renderer_di_bindings(r, C.byref(d)), "di bindings")
            dres = download_di(d.reservoirPrev)  # after the flip: the last frame's
        # after the flip: grisReservoirPrev holds the last frame's reservoirs
        return download(f.grisReservoirPrev), download(f.grisReservoir), out, dout, dres
    finally:
        Hh.rt_renderer_destroy(r)


def pipelined(Hh, p, hs, frames, ta, tb):
    for k in range(frames):
        rt.check(Hh.rt_pipeline_frame(p, hs.camera(k, W, H), ta, tb), "pipeline frame")
    f = rt.Frame()
    rt.check(Hh.rt_pipeline_bindings(p, C.byref(f)), "bindings")
    out = np.zeros((H, W, 4), np.float32)
    rt.check(Hh.rt_pipeline_read_output(p, 1, out.ctypes.data, tb), "read")
    return download(f.grisReservoir), download(f.grisReservoirPrev), out


@pytest.mark.parametrize("unbiased", [False, True])
def test_pipeline_matches_sequential_frames(unbiased):
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=6, cap=20, neighbors=3, radius=20.0,
                    unbiased=unbiased)
    print(f"[gpu] pipeline vs sequential, unbiased={unbiased}", flush=True)
    want_last, want_prev, want_out, _, _ = sequential(hs, ds, s, FRAMES)
    assert (want_last["F"] != 0).any()
    Hh = rt.host()
    p = C.c_void_p()
    rt.check(Hh.rt_pipeline_create(ds.handle, W, H, C.byref(p)), "pipeline")
    try:
        rt.check(Hh.rt_pipeline_set_settings(p, C.byref(s)), "settings")
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        ta, tb = (C.c_void_p(x.cuda_stream) for x in streams)
        torch.cuda.synchronize()
        for run in range(2):  # the second walk after a reset must repeat the first
            last, prev, out = pipelined(Hh, p, hs, FRAMES, ta, tb)
            assert not gpuhelp.compare_reservoirs(want_last, last).any(), f"walk {run}: last frame differs"
            assert not gpuhelp.compare_reservoirs(want_prev, prev).any(), f"walk {run}: frame before differs"
            assert gpuhelp.same_float(want_out, out).all(), f"walk {run}: output differs"
            rt.check(Hh.rt_pipeline_reset_reservoirs(p, gpuhelp.stream()), "reset")
        # one stream for both: the frames run one after the other and still agree
        last, prev, out = pipelined(Hh, p, hs, FRAMES, ta, ta)
        assert not gpuhelp.compare_reservoirs(want_last, last).any()
        assert gpuhelp.same_float(want_out, out).all()
        # a smaller path-trace launch (a few workgroups) changes nothing either
        rt.check(Hh.rt_pipeline_reset_reservoirs(p, gpuhelp.stream()), "reset")
        rt.check(Hh.rt_pipeline_set_trace_groups(p, 37), "trace groups")
        last, prev, out = pipelined(Hh, p, hs, FRAMES, ta, tb)
        assert not gpuhelp.compare_reservoirs(want_last, last).any()
        assert gpuhelp.same_float(want_out, out).all()
    finally:
        torch.cuda.synchronize()
        Hh.rt_pipeline_destroy(p)


@pytest.mark.parametrize("direct,third", [(1, False), (2, False), (2, True)],
                         ids=["naive", "restir_di", "restir_di_own_stream"])
def test_pipeline_direct_light_matches_sequential_frames(direct, third):
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=6, cap=20, neighbors=3, radius=20.0)
    di = rt.di_settings(shift=0, sample=2, temporal=True, spatial=True)
    print(f"[gpu] pipeline with direct method {direct} vs sequential", flush=True)
    want_last, _, want_out, want_direct, want_di = sequential(hs, ds, s, FRAMES, direct=direct, di=di)
    assert (want_direct[..., :3] > 0).any()
    Hh = rt.host()
    p = C.c_void_p()
    rt.check(Hh.rt_pipeline_create(ds.handle, W, H, C.byref(p)), "pipeline")
    try:
        rt.check(Hh.rt_pipeline_set_settings(p, C.byref(s)), "settings")
        rt.check(Hh.rt_pipeline_set_di_settings(p, C.byref(di)), "di settings")
        rt.check(Hh.rt_pipeline_set_direct_method(p, direct), "direct method")
        streams = [torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream()]
        ta, tb, tc = (C.c_void_p(x.cuda_stream) for x in streams)
        if third:
            rt.check(Hh.rt_pipeline_set_direct_stream(p, tc), "direct stream")
        torch.cuda.synchronize()
        last, _, out = pipelined(Hh, p, hs, FRAMES, ta, tb)
        direct_out = np.zeros((H, W, 4), np.float32)
        # the naive direct light is written on the trace stream, ReSTIR DI's on its own stream or the reuse one
        rt.check(Hh.rt_pipeline_read_output(p, 0, direct_out.ctypes.data, ta if direct == 1 else (tc if third else tb)),
                 "read direct")
        assert not gpuhelp.compare_reservoirs(want_last, last).any()
        assert gpuhelp.same_float(want_out, out).all()
        assert gpuhelp.same_float(want_direct, direct_out).all()
        if direct == 2:
            d = rt.DIReservoirs()
            rt.check(Hh.rt_pipeline_di_bindings(p, C.byref(d)), "di bindings")
            assert not gpuhelp.compare_di(want_di, download_di(d.reservoir)).any()
    finally:
        torch.cuda.synchronize()
        Hh.rt_pipeline_destroy(p)


def test_pipeline_refuses_learned_forms_and_wrong_film_size():
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    Hh = rt.host()
    p = C.c_void_p()
    rt.check(Hh.rt_pipeline_create(ds.handle, W, H, C.byref(p)), "pipeline")
    try:
        st = gpuhelp.stream()
        for kw in ({"learned": True}, {"guided": True}):
            s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=4, **kw)
            assert Hh.rt_pipeline_set_settings(p, C.byref(s)) == INVALID_ARGUMENT
        s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=4)
        rt.check(Hh.rt_pipeline_set_settings(p, C.byref(s)), "settings")
        assert Hh.rt_pipeline_frame(p, hs.camera(0, W + 8, H), st, st) == INVALID_ARGUMENT
        assert Hh.rt_pipeline_set_direct_method(p, 3) == INVALID_ARGUMENT
        rt.check(Hh.rt_pipeline_frame(p, hs.camera(0, W, H), st, st), "frame")
    finally:
        torch.cuda.synchronize()
        Hh.rt_pipeline_destroy(p)
