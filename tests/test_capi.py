"""The C-ABI libraries load and export every function declared in include/*.h; host-side
argument checks reject bad input without touching a GPU."""
import ctypes as C
import os
import re

import pytest

import rtpkg

rt = rtpkg.rt
INCLUDE = os.path.join(rtpkg.ROOT, "include")


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("restirpt.h", rt.core), ("restirpt_host.h", rt.host)])
def test_exports_every_declared_symbol(header, loader):
    lib = loader()
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_python_binding_covers_the_abi():
    assert set(declared("restirpt.h")) == set(rt.CORE_SYMBOLS)
    assert set(declared("restirpt_host.h")) == set(rt.HOST_SYMBOLS)


def test_struct_sizes_match_reference_layouts():
    assert C.sizeof(rt.Camera) == 352          # layouts.glsl Camera (std140 uniform)
    assert C.sizeof(rt.GrisReservoir) == 96    # layouts.glsl GRISReservoir (std430)
    assert C.sizeof(rt.ObjectInstance) == 224
    assert C.sizeof(rt.TriangleLight) == 64
    assert C.sizeof(rt.Material) == 32 and C.sizeof(rt.MeshVertex) == 32
    # 20-B push constant + 3 kernel constants + target function + resampling-weight mode + proposal
    assert C.sizeof(rt.GrisSettings) == 44
    assert C.sizeof(rt.MlpConfig) == 72
    # rt_view: scene pointer + 2 cameras + rt_frame (112 B) + settings (44 B) + 4 B padding + 2 networks
    assert C.sizeof(rt.Frame) == 112 and C.sizeof(rt.View) == 8 + 2 * 352 + 112 + 44 + 4 + 2 * 8
    assert C.sizeof(rt.BatchTuning) == 24
    assert C.sizeof(rt.DIReservoir) == 64      # layouts.glsl DIReservoir (std430)
    assert C.sizeof(rt.DISettings) == 16       # TestReSTIR::Settings push constant
    assert C.sizeof(rt.DIReservoirs) == 24
    assert C.sizeof(rt.PostSettings) == 16     # PostProcessFrag::PushConstant
    assert C.sizeof(rt.GIReservoirs) == 16
    assert rt.core().rt_gi_reservoir_bytes(1920, 1080) == 1920 * 1080 * 48
    assert rt.core().rt_di_reservoir_bytes(1920, 1080) == 1920 * 1080 * 64
    # five 16-B planes of reference fields + the learned-target plane = the reference's 96-B reservoir
    assert rt.core().rt_gris_reservoir_bytes(1920, 1080) == 1920 * 1080 * 96


def test_argument_checks_fail_loudly():
    L = rt.core()
    cams = (rt.Camera * 2)()
    f = rt.Frame()
    s = rt.settings()
    assert L.rt_gris_path_trace(None, cams, C.byref(f), C.byref(s), None) == 1
    assert b"scene is null" in L.rt_last_error()
    # a scene whose light table does not hold numLights + 1 entries is rejected before any upload
    a = rt.SceneArrays()
    mat = rt.Material()
    mat.textureIdx = 0xFFFFFFFF
    a.materials, a.numMaterials = C.addressof(mat), 1
    a.numLights, a.numLightTable = 3, 1
    tab = (rt.LightTableEntry * 1)()
    a.lightTable = C.addressof(tab)
    h = C.c_void_p()
    assert L.rt_device_scene_create(C.byref(a), None, C.byref(h)) == 1
    assert b"light table" in L.rt_last_error()
    # a material naming a texture the scene does not hold: rejected, not silently untextured
    mat.textureIdx = 0
    a.numLights, a.numLightTable = 0, 1
    tab[0].prob, tab[0].failId = 0.0, 0
    assert L.rt_device_scene_create(C.byref(a), None, C.byref(h)) == 1
    assert b"textureIdx out of range" in L.rt_last_error()
    bad = rt.settings(max_depth=0)
    fake = C.c_void_p(1)
    f.width, f.height = 4, 4
    cams[0].filmSize[0], cams[0].filmSize[1] = 4, 4
    f.depthNormal = f.albedoMatId = 8
    assert L.rt_gris_path_trace(fake, cams, C.byref(f), C.byref(bad), None) == 1
    assert b"maxDepth" in L.rt_last_error()
    r = rt.DIReservoirs()
    assert L.rt_di_render(fake, cams, C.byref(f), C.byref(r), C.byref(rt.di_settings(shift=3)), None) == 1
    assert b"shiftType" in L.rt_last_error()
    assert L.rt_di_render(fake, cams, C.byref(f), C.byref(r), C.byref(rt.di_settings(sample=3)), None) == 1
    assert b"sampleType" in L.rt_last_error()
    assert L.rt_di_path_gen(fake, cams, C.byref(f), None, C.byref(rt.di_settings()), None) == 1
    assert b"reservoirs is null" in L.rt_last_error()
    # frame graphs: every form's requirements are checked before anything is allocated
    fs = (rt.Frame * 2)()
    for j in range(2):
        fs[j].width, fs[j].height = 4, 4
    g = C.c_void_p()
    guided = rt.settings(guided=True)
    assert L.rt_frame_graph_create(fake, fs, C.byref(guided), None, None, C.byref(g)) == 1
    assert b"proposal network" in L.rt_last_error()
    wide = rt.settings(unbiased=True, neighbors=9)
    assert L.rt_frame_graph_create(fake, fs, C.byref(wide), None, None, C.byref(g)) == 1
    assert b"at most 8" in L.rt_last_error()
    learned = rt.settings(learned=True)
    assert L.rt_frame_graph_create(fake, fs, C.byref(learned), None, None, C.byref(g)) == 1
    assert b"radiance network" in L.rt_last_error()
    # the graph always records temporal reuse: both parities need the previous G-buffer
    plain = rt.settings()
    for j in range(2):
        for name in ("depthNormal", "albedoMatId", "motionVector", "grisReservoir", "grisReservoirPrev",
                     "grisReservoirTemp", "indirectOutput"):
            setattr(fs[j], name, 8)
    assert L.rt_frame_graph_create(fake, fs, C.byref(plain), None, None, C.byref(g)) == 1
    assert b"previous G-buffer" in L.rt_last_error()
    for j in range(2):
        fs[j].depthNormalPrev = fs[j].albedoMatIdPrev = 8
    fs[1].albedoMatIdPrev = None
    assert L.rt_frame_graph_create(fake, fs, C.byref(plain), None, None, C.byref(g)) == 1
    assert b"previous G-buffer" in L.rt_last_error()
    fs[1].albedoMatIdPrev = 8
    # a scene without lights is rejected as the eager passes reject it (a zero-filled handle has none)
    unlit = C.create_string_buffer(1 << 16)
    assert L.rt_frame_graph_create(unlit, fs, C.byref(plain), None, None, C.byref(g)) == 1
    assert b"no lights" in L.rt_last_error()
    # eager entry points: the learned refresh takes a radiance network only, and the one-call
    # frame refuses the learned proposal, whose refresh and training it does not run
    f.grisReservoir = 8
    not_radiance = C.create_string_buffer(1 << 16)  # a zero-filled network handle has 0 outputs
    assert L.rt_gris_learned_refresh(fake, not_radiance, cams, C.byref(f), None) == 1
    assert b"not a radiance network" in L.rt_last_error()
    assert L.rt_gris_render(fake, cams, C.byref(f), C.byref(guided), None) == 1
    assert b"rt_gris_guide_refresh" in L.rt_last_error()
    # the capped path trace checks its arguments as rt_gris_path_trace does
    assert L.rt_gris_path_trace_groups(None, cams, C.byref(f), C.byref(s), 64, None) == 1
    assert b"scene is null" in L.rt_last_error()
    assert L.rt_gris_path_trace_groups(fake, cams, C.byref(f), C.byref(bad), 64, None) == 1
    assert b"maxDepth" in L.rt_last_error()
    # frame pipelines: null handles and arguments are refused before any device work
    Hh = rt.host()
    pipe = C.c_void_p()
    assert Hh.rt_pipeline_create(None, 4, 4, C.byref(pipe)) == 1
    assert Hh.rt_pipeline_create(fake, 0, 4, C.byref(pipe)) == 1
    assert Hh.rt_pipeline_destroy(None) == 1
    assert Hh.rt_pipeline_set_settings(None, C.byref(s)) == 1
    assert Hh.rt_pipeline_set_trace_groups(None, 64) == 1
    assert Hh.rt_pipeline_frame(None, cams, None, None) == 1
    assert Hh.rt_pipeline_bindings(None, C.byref(f)) == 1
    assert Hh.rt_pipeline_read_output(None, 1, None, None) == 1
    assert Hh.rt_pipeline_set_direct_method(None, 1) == 1
    assert Hh.rt_pipeline_set_di_settings(None, C.byref(rt.di_settings())) == 1
    assert Hh.rt_pipeline_di_bindings(None, C.byref(rt.DIReservoirs())) == 1
    assert Hh.rt_pipeline_set_direct_stream(None, None) == 1
    assert Hh.rt_pipeline_reset_reservoirs(None, None) == 1


def test_missing_library_is_an_error(monkeypatch):
    monkeypatch.setattr(rt, "_core", None)
    monkeypatch.setattr(rt, "CORE_LIB", "/nonexistent/librestirpt.so")
    with pytest.raises(RuntimeError, match="missing"):
        rt.core()
