"""GPU parity of ReSTIR GI temporal reuse (gi_resample_temporal) against the oracle, through the C ABI:
the reservoirs bit for bit and the accumulated indirect output, over frame sequences with camera motion
on three scenes (including the 1.1M-triangle terrain); the renderer's indirect method 2; the layout."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt
GI_FLOAT = ["bary", "rcLo", "resampleWeight", "contribWeight", "pad0"]
GI_UINT = ["instanceIdx", "triangleIdx", "rcPrevCoord", "sampleCount"]


def compare_gi(a, b):
    bad = np.zeros(a.shape[0], bool)
    for f in GI_FLOAT:
        x, y = a[f].reshape(a.shape[0], -1), b[f].reshape(a.shape[0], -1)
        bad |= ~((x == y) | (np.isnan(x) & np.isnan(y))).all(axis=1)
    for f in GI_UINT:
        bad |= a[f] != b[f]
    return bad


class GIRun(gpuhelp.GrisRun):
    def __init__(self, ds, W, H):
        super().__init__(ds, W, H)
        rb = rt.core().rt_gi_reservoir_bytes(W, H)
        self.gi = [gpuhelp.dev_empty(rb), gpuhelp.dev_empty(rb)]
        for b in self.gi:
            rt.check(rt.core().rt_gi_reservoir_reset(b.data_ptr(), W, H, gpuhelp.stream()), "gi reset")

    def gi_temporal(self, cams):
        f = self.frame_struct()
        r = rt.GIReservoirs()
        r.reservoir, r.reservoirPrev = self.gi[self.cur].data_ptr(), self.gi[self.cur ^ 1].data_ptr()
        rt.check(rt.core().rt_gi_temporal(self.ds.handle, cams, C.byref(f), C.byref(r), gpuhelp.stream()), "gi")

    def gi_reservoirs(self, buf):
        out = np.zeros(self.W * self.H, orc.GI_DTYPE)
        torch.cuda.synchronize()
        rt.check(rt.core().rt_gi_reservoir_download(buf.data_ptr(), self.W, self.H, out.ctypes.data, gpuhelp.stream()),
                 "download")
        return out


@pytest.mark.parametrize("config,W,H,frames", [(0, 72, 45, 5), (2, 40, 32, 3), (1, 64, 40, 3)],
                         ids=["boxroom", "floor", "terrain"])
def test_gi_sequence_parity(config, W, H, frames):
    hs = rt.HostScene(config, 0)
    ds = rt.DeviceScene(hs.arrays)
    osc = orc.OracleScene(hs.arrays)
    run = GIRun(ds, W, H)
    n = W * H
    o_resv = [orc.reset_gi_reservoirs(n), orc.reset_gi_reservoirs(n)]
    o_out = np.zeros((H, W, 4), np.float32)
    o_dn, o_am = [None, None], [None, None]
    cur = 0
    for frame in range(frames):
        print(f"[gpu] gi config {config} {W}x{H} frame {frame}", flush=True)
        cams = hs.camera(frame, W, H)
        run.gbuffer(cams[0])
        dn, am, mv, _ = osc.gbuffer(cams[0])
        o_dn[cur], o_am[cur] = dn, am
        prev = cur ^ 1
        pdn = o_dn[prev] if o_dn[prev] is not None else np.zeros_like(dn)
        pam = o_am[prev] if o_am[prev] is not None else np.zeros_like(am)
        run.gi_temporal(cams)
        osc.gi_temporal(cams, dn, am, mv, pdn, pam, o_resv[cur], o_resv[prev], o_out)
        bad = compare_gi(run.gi_reservoirs(run.gi[run.cur]), o_resv[cur])
        assert not bad.any(), f"frame {frame}: {int(bad.sum())} reservoirs differ"
        out = run.output(1)
        assert np.array_equal(out.view(np.uint32), o_out.view(np.uint32)), \
            f"frame {frame}: {int((out != o_out).any(-1).sum())} output pixels differ"
        run.flip()
        cur ^= 1
    assert np.isfinite(o_out).all()
    if config != 2:
        assert o_out[..., :3].mean() > 0


def test_renderer_gi_method_and_layout():
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 48, 32
    Hh = rt.host()
    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
    try:
        run = GIRun(ds, W, H)
        for frame in range(3):
            cams = hs.camera(frame, W, H)
            rt.check(Hh.rt_renderer_frame_methods(r, cams, 0, 2, gpuhelp.stream()), "frame")
            run.gbuffer(cams[0])
            run.gi_temporal(cams)
            torch.cuda.synchronize()
            host = np.zeros((H, W, 4), np.float32)
            rt.check(Hh.rt_renderer_read_output(r, 1, host.ctypes.data, gpuhelp.stream()), "read")
            assert np.array_equal(host.view(np.uint32), run.output(1).view(np.uint32)), frame
            run.flip()
    finally:
        Hh.rt_renderer_destroy(r)
    # layout round trip: plane k = bytes 16k..16k+15
    n = 35
    g = np.random.default_rng(5)
    host = np.frombuffer(g.integers(0, 2 ** 32, size=n * 12, dtype=np.uint32).tobytes(), orc.GI_DTYPE).copy()
    buf = gpuhelp.dev_empty(rt.core().rt_gi_reservoir_bytes(7, 5))
    rt.check(rt.core().rt_gi_reservoir_upload(host.ctypes.data, 7, 5, buf.data_ptr(), gpuhelp.stream()), "up")
    back = np.zeros(n, orc.GI_DTYPE)
    rt.check(rt.core().rt_gi_reservoir_download(buf.data_ptr(), 7, 5, back.ctypes.data, gpuhelp.stream()), "down")
    assert back.tobytes() == host.tobytes()
    planes = gpuhelp.to_host(buf, np.uint32, (3, n, 4))
    assert np.array_equal(planes, host.view(np.uint32).reshape(n, 3, 4).transpose(1, 0, 2))
