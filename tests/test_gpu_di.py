"""GPU parity of ReSTIR DI (the TestReSTIR passes) against the oracle, through the C ABI: every
pass's reservoirs bit for bit and the accumulated direct output, over frame sequences with camera
motion, for each shift mapping and sample type; plus the reservoir layout round trip and the
argument checks."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt


@pytest.fixture(scope="module", params=[0, 2], ids=["boxroom", "floor"])
def scene(request):
    print(f"[gpu] di scene config {request.param}", flush=True)
    hs = rt.HostScene(request.param, 0)
    return hs, rt.DeviceScene(hs.arrays), orc.OracleScene(hs.arrays)


def run_sequence(hs, ds, osc, W, H, s, frames, render=False):
    """frames of path gen -> temporal -> spatial on GPU and oracle, compared after every pass."""
    run = gpuhelp.DIRun(ds, W, H)
    n = W * H
    o_dn, o_am = [None, None], [None, None]
    o_resv = [orc.reset_di_reservoirs(n), orc.reset_di_reservoirs(n)]
    o_temp = orc.reset_di_reservoirs(n)
    o_out = np.zeros((H, W, 4), np.float32)
    cur = 0
    for frame in range(frames):
        print(f"[gpu] di shift {s.shiftType} sample {s.sampleType} {W}x{H} frame {frame}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        osc.di_path_gen(cams, dn, am, s, o_resv[cur])
        osc.di_temporal(cams, dn, am, mv, pdn, pam, s, o_resv[cur], o_resv[prev], o_temp)
        if render:
            run.di_render(cams, s)
        else:
            run.di_path_gen(cams, s)
            bad = gpuhelp.compare_di(run.di_reservoirs(run.di[run.cur]), o_resv[cur])
            assert not bad.any(), f"frame {frame} path gen: {int(bad.sum())} reservoirs differ"
            run.di_temporal(cams, s)
        bad = gpuhelp.compare_di(run.di_reservoirs(run.di_temp), o_temp)
        assert not bad.any(), f"frame {frame} temporal: {int(bad.sum())} reservoirs differ"
        if not render:
            run.di_spatial(cams, s)
        osc.di_spatial(cams, dn, am, s, o_temp, o_resv[cur], o_out)
        bad = gpuhelp.compare_di(run.di_reservoirs(run.di[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame} spatial: {int(bad.sum())} reservoirs differ"
        out = run.output(0)
        assert np.array_equal(out.view(np.uint32), o_out.view(np.uint32)), \
            f"frame {frame}: {int((out != o_out).any(-1).sum())} direct-output pixels differ"
        run.flip()
        cur ^= 1
    return o_out, o_resv[cur ^ 1]


@pytest.mark.parametrize("shift,sample", [(0, 0), (0, 2), (1, 0), (1, 2), (0, 1), (2, 2)],
                         ids=["rc-light", "rc-both", "replay-light", "replay-both", "rc-bsdf", "hybrid-both"])
def test_di_sequence_parity(scene, shift, sample):
    hs, ds, osc = scene
    s = rt.di_settings(shift=shift, sample=sample, temporal=True, spatial=True)
    out, resv = run_sequence(hs, ds, osc, 72, 45, s, 4)
    assert np.isfinite(out).all() and (out[..., :3] >= 0).all()
    if sample != 1:
        assert out[..., :3].mean() > 0


def test_di_render_and_reuse_switches(scene):
    """rt_di_render = the three passes in order; reuse switched off pass by pass."""
    hs, ds, osc = scene
    for temporal, spatial in ((False, True), (True, False), (False, False)):
        s = rt.di_settings(shift=0, sample=2, temporal=temporal, spatial=spatial)
        run_sequence(hs, ds, osc, 40, 24, s, 2, render=True)


def test_di_terrain_parity():
    """The 1.1M-triangle terrain (the benchmark scene) at a small size, default settings."""
    hs = rt.HostScene(1, 0)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    s = rt.di_settings(shift=0, sample=0, temporal=True, spatial=True)
    out, _ = run_sequence(hs, ds, osc, 64, 40, s, 3)
    assert out[..., :3].mean() > 0


def test_di_reservoir_layout_round_trip():
    W, H = 13, 7
    n = W * H
    g = np.random.default_rng(3)
    host = np.frombuffer(g.integers(0, 2 ** 32, size=n * 16, dtype=np.uint32).tobytes(), orc.DI_DTYPE).copy()
    buf = gpuhelp.dev_empty(rt.core().rt_di_reservoir_bytes(W, H))
    L = rt.core()
    rt.check(L.rt_di_reservoir_upload(host.ctypes.data, W, H, buf.data_ptr(), gpuhelp.stream()), "upload")
    back = np.zeros(n, orc.DI_DTYPE)
    rt.check(L.rt_di_reservoir_download(buf.data_ptr(), W, H, back.ctypes.data, gpuhelp.stream()), "download")
    assert back.tobytes() == host.tobytes()
    # the device layout is four planes, plane k = bytes 16k..16k+15 of each reservoir
    planes = gpuhelp.to_host(buf, np.uint32, (4, n, 4))
    raw = host.view(np.uint32).reshape(n, 4, 4)
    assert np.array_equal(planes, raw.transpose(1, 0, 2))
    rt.check(L.rt_di_reservoir_reset(buf.data_ptr(), W, H, gpuhelp.stream()), "reset")
    rt.check(L.rt_di_reservoir_download(buf.data_ptr(), W, H, back.ctypes.data, gpuhelp.stream()), "download")
    assert back.tobytes() == orc.reset_di_reservoirs(n).tobytes()


def test_di_stale_reservoirs_stay_in_bounds(scene):
    """Reservoir memory holding random bits (never cleared) must not send a kernel outside the scene
    arrays: sample references are range-checked. The passes run and the output stays finite."""
    hs, ds, osc = scene
    W, H = 48, 32
    n = W * H
    run = gpuhelp.DIRun(ds, W, H)
    g = np.random.default_rng(11)
    junk = np.frombuffer(g.integers(0, 2 ** 32, size=n * 16, dtype=np.uint32).tobytes(), orc.DI_DTYPE).copy()
    junk["isLightSample"] &= 1
    L = rt.core()
    for b in run.di + [run.di_temp]:
        rt.check(L.rt_di_reservoir_upload(junk.ctypes.data, W, H, b.data_ptr(), gpuhelp.stream()), "upload")
    s = rt.di_settings(shift=0, sample=2, temporal=True, spatial=True)
    for frame in (1, 2):
        print(f"[gpu] stale reservoirs frame {frame}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        run.di_render(cams, s)
        run.flip()
    torch.cuda.synchronize()
    out = run.output(0)
    assert not np.isnan(out).any()


def test_di_argument_checks():
    L = rt.core()
    cams = (rt.Camera * 2)()
    f = rt.Frame()
    r = rt.DIReservoirs()
    s = rt.di_settings()
    assert L.rt_di_path_gen(None, cams, C.byref(f), C.byref(r), C.byref(s), None) == 1
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    run = gpuhelp.DIRun(ds, 16, 8)
    cams = hs.camera(0, 16, 8)
    f = run.frame_struct()
    r = run.di_struct()
    bad = rt.di_settings(shift=3)
    assert L.rt_di_render(ds.handle, cams, C.byref(f), C.byref(r), C.byref(bad), None) == 1
    assert b"shiftType" in L.rt_last_error()
    bad = rt.di_settings(sample=5)
    assert L.rt_di_render(ds.handle, cams, C.byref(f), C.byref(r), C.byref(bad), None) == 1
    assert b"sampleType" in L.rt_last_error()
    r.reservoirTemp = None
    assert L.rt_di_temporal(ds.handle, cams, C.byref(f), C.byref(r), C.byref(s), None) == 1


def test_renderer_methods_match_pass_calls():
    """rt_renderer_frame_methods(direct = ReSTIR DI, indirect = ReSTIR PT) equals the G-buffer pass,
    rt_di_render and rt_gris_render called in that order on separate buffers, frame after frame."""
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 56, 40
    Hh = rt.host()
    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
    try:
        gs = rt.settings(shift=2, temporal=True, spatial=True, max_depth=5)
        rt.check(Hh.rt_renderer_set_settings(r, C.byref(gs)), "settings")
        di = rt.di_settings(shift=1, sample=2, temporal=True, spatial=True)
        rt.check(Hh.rt_renderer_set_di_settings(r, C.byref(di)), "di settings")
        run = gpuhelp.DIRun(ds, W, H)
        for frame in range(3):
            print(f"[gpu] renderer methods frame {frame}", flush=True)
            cams = hs.camera(frame, W, H)
            rt.check(Hh.rt_renderer_frame_methods(r, cams, 2, 3, gpuhelp.stream()), "frame")
            run.gbuffer(cams[0])
            run.di_render(cams, di)
            f = run.frame_struct()
            rt.check(rt.core().rt_gris_render(ds.handle, cams, C.byref(f), C.byref(gs), gpuhelp.stream()), "gris")
            torch.cuda.synchronize()
            for which in (0, 1):
                host = np.zeros((H, W, 4), np.float32)
                rt.check(Hh.rt_renderer_read_output(r, which, host.ctypes.data, gpuhelp.stream()), "read")
                assert np.array_equal(host.view(np.uint32), run.output(which).view(np.uint32)), (frame, which)
            run.flip()
        assert run.output(0)[..., :3].mean() > 0 and run.output(1)[..., :3].mean() > 0
        assert Hh.rt_renderer_frame_methods(r, cams, 3, 3, gpuhelp.stream()) != 0
    finally:
        Hh.rt_renderer_destroy(r)
