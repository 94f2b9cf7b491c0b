"""GPU parity on a scene loaded from a file in the reference's format (rt_scene_load of
golden/scenes/room.xml: OBJ meshes with transforms, MTL colours, metal and MetallicWorkflow
overrides, smooth normals, an emitter of given power), against the oracle fed the same arrays:
LBVH, G-buffer, plain PT, three GRIS frames per shift mapping and three ReSTIR DI frames, pass by
pass; and the renderer's screenshot of the file's view at the file's film size."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402
import rtpkg  # noqa: E402
from test_gpu_di import run_sequence as di_sequence  # noqa: E402
from test_gpu_parity import close, run_gris_frames  # noqa: E402
from test_post import read_png  # noqa: E402

rt = gpuhelp.rt
SCENE = Path(rtpkg.ROOT) / "tests" / "golden" / "scenes" / "room.xml"


@pytest.fixture(scope="module")
def room():
    print(f"[gpu] scene file {SCENE.name}", flush=True)
    hs = rt.HostScene.load(SCENE)
    W, H = hs.film_size()
    return hs, rt.DeviceScene(hs.arrays), orc.OracleScene(hs.arrays), W, H


def test_lbvh_and_gbuffer_bitexact(room):
    hs, ds, osc, W, H = room
    codes, order, nodes, info = gpuhelp.gpu_bvh(ds)
    ocodes, oorder, onodes, obox = osc.bvh()
    assert info.numPrimitives == osc.n == 44
    assert np.array_equal(codes, ocodes) and np.array_equal(order, oorder) and np.array_equal(nodes, onodes)
    assert np.array_equal(np.array(info.sceneLo, np.float32), obox[:3])
    assert np.array_equal(gpuhelp.gpu_bvh4(ds), osc.bvh4())
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    dn, am, mv = run.gbuffer_host()
    odn, oam, omv, _ = osc.gbuffer(cams[0])
    assert np.array_equal(dn.view(np.uint32), odn.view(np.uint32))
    assert np.array_equal(am, oam) and np.array_equal(mv, omv)
    assert len(np.unique(am[..., 1] >> 16)) == 5  # white, red, green, metal ball, box


def test_pt_plain_parity(room):
    hs, ds, osc, W, H = room
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    run.plain(cams, 5)
    dn, am, _ = run.gbuffer_host()
    od, oi, _ = osc.pt_plain(cams[0], dn, am, 5)
    for name, g, o in (("direct", run.output(0), od), ("indirect", run.output(1), oi)):
        ok = close(g, o).all(axis=-1)
        assert ok.all(), f"{name}: {int((~ok).sum())} pixels differ"
    assert run.output(0)[..., :3].mean() > 0 and run.output(1)[..., :3].mean() > 0


@pytest.mark.parametrize("shift", [2, 0], ids=["hybrid", "reconnection"])
def test_gris_sequence_parity(room, shift):
    hs, ds, osc, W, H = room
    s = rt.settings(shift=shift, temporal=True, spatial=True, max_depth=6, cap=20, radius=20.0)
    report = run_gris_frames(hs, ds, osc, W, H, s, 3)
    assert all(r == (0, 0, 0, 0, 0) for r in report), report


def test_di_sequence_parity(room):
    hs, ds, osc, W, H = room
    out, _ = di_sequence(hs, ds, osc, W, H, rt.di_settings(shift=0, sample=2, temporal=True, spatial=True), 3)
    assert np.isfinite(out).all() and out[..., :3].mean() > 0


def test_renderer_screenshot_of_file_view(room, tmp_path):
    hs, ds, _, W, H = room
    Hh = rt.host()
    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "renderer")
    try:
        for frame in range(3):
            rt.check(Hh.rt_renderer_frame_methods(r, hs.camera(frame, W, H), 2, 3, gpuhelp.stream()), "frame")
        path = str(tmp_path / "room.png")
        s = rt.PostSettings(1, 1, 0, 0)
        rt.check(Hh.rt_renderer_screenshot(r, C.byref(s), path.encode(), gpuhelp.stream()), "screenshot")
        outs = []
        for which in (0, 1):
            host = np.zeros((H, W, 4), np.float32)
            rt.check(Hh.rt_renderer_read_output(r, which, host.ctypes.data, gpuhelp.stream()), "read")
            outs.append(host)
        img = read_png(path)
        assert img.shape == (H, W, 4)
        assert np.array_equal(img, orc.post_process(outs[0], outs[1], (1, 1, 0, 0)))
        # the red wall is red and the green wall green in the picture
        left, right = img[H // 2, 2, :3].astype(int), img[H // 2, W - 3, :3].astype(int)
        assert left[0] > left[1] and right[1] > right[0]
    finally:
        Hh.rt_renderer_destroy(r)
