"""GPU post-processing (rt_post_process) against the oracle, byte for byte, for every tone curve,
gamma and output switch; and the renderer's screenshot end to end (render, post-process, PNG)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402
from test_post import read_png  # noqa: E402

rt = gpuhelp.rt


@pytest.mark.parametrize("W,H", [(37, 23), (256, 144)])
def test_post_process_matches_oracle(W, H):
    g = np.random.default_rng(W)
    direct = np.exp(g.uniform(-6, 3, size=(H, W, 4))).astype(np.float32)
    indirect = np.exp(g.uniform(-6, 3, size=(H, W, 4))).astype(np.float32)
    direct[0, 0, :3] = 0.0
    d_dir, d_ind = torch.from_numpy(direct).cuda(), torch.from_numpy(indirect).cuda()
    out = torch.zeros(W * H * 4, dtype=torch.uint8, device="cuda")
    f = rt.Frame()
    f.width, f.height = W, H
    f.directOutput, f.indirectOutput = d_dir.data_ptr(), d_ind.data_ptr()
    for tone in (0, 1, 2):
        for gamma in (0, 1):
            for nd, ni in ((0, 0), (1, 0), (0, 1)):
                print(f"[gpu] post {W}x{H} tone {tone} gamma {gamma} off {nd}{ni}", flush=True)
                s = rt.PostSettings(tone, gamma, nd, ni)
                rt.check(rt.core().rt_post_process(C.byref(f), C.byref(s), out.data_ptr(), gpuhelp.stream()), "post")
                torch.cuda.synchronize()
                got = out.cpu().numpy().reshape(H, W, 4)
                want = orc.post_process(direct, indirect, (tone, gamma, nd, ni))
                assert np.array_equal(got, want), f"{int((got != want).any(-1).sum())} pixels differ"
    bad = rt.PostSettings(3, 1, 0, 0)
    assert rt.core().rt_post_process(C.byref(f), C.byref(bad), out.data_ptr(), gpuhelp.stream()) != 0


def test_renderer_screenshot(tmp_path):
    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 64, 48
    Hh = rt.host()
    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
    try:
        for frame in range(2):
            rt.check(Hh.rt_renderer_frame_methods(r, hs.camera(frame, W, H), 2, 3, gpuhelp.stream()), "frame")
        path = str(tmp_path / "shot.png")
        s = rt.PostSettings(1, 1, 0, 0)  # the Renderer's defaults: Filmic + gamma
        rt.check(Hh.rt_renderer_screenshot(r, C.byref(s), path.encode(), gpuhelp.stream()), "screenshot")
        outs = []
        for which in (0, 1):
            host = np.zeros((H, W, 4), np.float32)
            rt.check(Hh.rt_renderer_read_output(r, which, host.ctypes.data, gpuhelp.stream()), "read")
            outs.append(host)
        img = read_png(path)
        assert np.array_equal(img, orc.post_process(outs[0], outs[1], (1, 1, 0, 0)))
        assert img[..., :3].mean() > 5
    finally:
        Hh.rt_renderer_destroy(r)
