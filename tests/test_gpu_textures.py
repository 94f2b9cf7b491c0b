"""GPU parity on a textured scene file (a nearest-filtered checker through <baseColor image=...>, a
linearly filtered MTL map_Kd tiled four times over the floor; tests/test_textures.py writes the
files), against the oracle fed the same arrays: the G-buffer's albedo (GBuffer.frag:33-37), plain
PT, three GRIS frames and three ReSTIR DI frames pass by pass (every surface albedo through
loadSurfaceInfo, ray_layouts.glsl:80-86), bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402
from test_gpu_di import run_sequence as di_sequence  # noqa: E402
from test_gpu_parity import close, run_gris_frames  # noqa: E402
from test_textures import write_textured_scene  # noqa: E402

rt = gpuhelp.rt


@pytest.fixture(scope="module")
def textured(tmp_path_factory):
    path, _, _ = write_textured_scene(tmp_path_factory.mktemp("textured"))
    print(f"[gpu] textured scene {path}", flush=True)
    hs = rt.HostScene.load(path)
    W, H = hs.film_size()
    return hs, rt.DeviceScene(hs.arrays), orc.OracleScene(hs.arrays), W, H


def test_textured_gbuffer_bitexact(textured):
    hs, ds, osc, W, H = textured
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    dn, am, mv = run.gbuffer_host()
    odn, oam, omv, _ = osc.gbuffer(cams[0])
    assert np.array_equal(dn.view(np.uint32), odn.view(np.uint32))
    assert np.array_equal(am, oam) and np.array_equal(mv, omv)
    # both textures show: the checker's two colours and a range of floor colours
    assert len(np.unique(am[..., 0][dn[..., 0] > 0])) > 6


def test_textured_pt_plain_parity(textured):
    hs, ds, osc, W, H = textured
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    run.plain(cams, 5)
    dn, am, _ = run.gbuffer_host()
    od, oi, _ = osc.pt_plain(cams[0], dn, am, 5)
    for name, g, o in (("direct", run.output(0), od), ("indirect", run.output(1), oi)):
        ok = close(g, o).all(axis=-1)
        assert ok.all(), f"{name}: {int((~ok).sum())} pixels differ"
    assert run.output(1)[..., :3].mean() > 0


def test_textured_gris_and_di_parity(textured):
    hs, ds, osc, W, H = textured
    s = rt.settings(shift=2, temporal=True, spatial=True, max_depth=6, cap=20, radius=20.0)
    report = run_gris_frames(hs, ds, osc, W, H, s, 3)
    assert all(r == (0, 0, 0, 0, 0) for r in report), report
    out, _ = di_sequence(hs, ds, osc, W, H, rt.di_settings(shift=0, sample=2, temporal=True, spatial=True), 3)
    assert np.isfinite(out).all() and out[..., :3].mean() > 0
