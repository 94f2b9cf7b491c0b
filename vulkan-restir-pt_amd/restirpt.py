"""Python bindings (ctypes) of the MI355X ReSTIR PT core.

Mirrors include/restirpt.h and include/restirpt_host.h. The structures below
have the byte layouts of the reference's GLSL structs (src/shader/layouts.glsl)
so numpy arrays can be handed to the C ABI without conversion.

torch is imported before the HIP libraries are loaded: torch ships its own
libamdhip64 with the same SONAME, and loading ours first would put two HIP
runtimes in one process. Device memory for tests and benchmarks is torch
tensors; the kernels themselves are in librestirpt.so and fail loudly when it
is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

try:  # one HIP runtime per process: torch's, when torch is present
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
CORE_LIB = os.path.join(HERE, "librestirpt.so")
HOST_LIB = os.path.join(HERE, "librestirpt_host.so")

RT_OK = 0

# ---------------------------------------------------------------------------------------
# Structures (include/restirpt.h)


class Material(C.Structure):
    _fields_ = [("baseColor", C.c_float * 3), ("type", C.c_uint32), ("textureIdx", C.c_uint32),
                ("metallic", C.c_float), ("roughness", C.c_float), ("ior", C.c_float)]


class MeshVertex(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("uvx", C.c_float), ("norm", C.c_float * 3), ("uvy", C.c_float)]


class ObjectInstance(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("transformInv", C.c_float * 16), ("transformInvT", C.c_float * 16),
                ("radiance", C.c_float * 3), ("pad0", C.c_float), ("indexOffset", C.c_uint32),
                ("indexCount", C.c_uint32), ("matIndex", C.c_uint32), ("pad2", C.c_float)]


class TriangleLight(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("nx", C.c_float), ("v1", C.c_float * 3), ("ny", C.c_float),
                ("v2", C.c_float * 3), ("nz", C.c_float), ("radiance", C.c_float * 3), ("area", C.c_float)]


class LightTableEntry(C.Structure):
    _fields_ = [("prob", C.c_float), ("failId", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("projView", C.c_float * 16),
                ("lastProjView", C.c_float * 16), ("pos", C.c_float * 3), ("FOV", C.c_float),
                ("angle", C.c_float * 3), ("nearZ", C.c_float), ("front", C.c_float * 3), ("farZ", C.c_float),
                ("right", C.c_float * 3), ("lensRadius", C.c_float), ("up", C.c_float * 3),
                ("focalDist", C.c_float), ("filmSize", C.c_uint32 * 2), ("frameIndex", C.c_uint32),
                ("seed", C.c_uint32)]


class GrisReservoir(C.Structure):
    _fields_ = [("bary", C.c_float * 2), ("instanceIdx", C.c_uint32), ("triangleIdx", C.c_uint32),
                ("rcLi", C.c_float * 3), ("rcRng", C.c_uint32), ("rcWi", C.c_float * 3), ("flags", C.c_uint32),
                ("pad", C.c_float * 2), ("rcPrevSamplePdf", C.c_float), ("rcJacobian", C.c_float),
                ("F", C.c_float * 3), ("primaryRng", C.c_uint32), ("sampleCount", C.c_float),
                ("resampleWeight", C.c_float), ("contribWeight", C.c_float), ("pad0", C.c_float)]


class DIReservoir(C.Structure):
    """DIReservoir (layouts.glsl:97-115), 64 B."""
    _fields_ = [("bary", C.c_float * 2), ("instanceIdx", C.c_uint32), ("triangleIdx", C.c_uint32),
                ("Li", C.c_float * 3), ("pad0", C.c_float), ("jacobian", C.c_float), ("samplePdf", C.c_float),
                ("rng", C.c_uint32), ("isLightSample", C.c_uint32), ("sampleCount", C.c_uint32),
                ("resampleWeight", C.c_float), ("contribWeight", C.c_float), ("weight", C.c_float)]


class DISettings(C.Structure):
    """TestReSTIR::Settings (src/TestReSTIR.h:12-17)."""
    _fields_ = [("shiftType", C.c_uint32), ("sampleType", C.c_uint32), ("temporalReuse", C.c_uint32),
                ("spatialReuse", C.c_uint32)]


class DIReservoirs(C.Structure):
    _fields_ = [("reservoir", C.c_void_p), ("reservoirPrev", C.c_void_p), ("reservoirTemp", C.c_void_p)]


class PostSettings(C.Structure):
    """PostProcessFrag::PushConstant; the Renderer's defaults are Filmic + gamma."""
    _fields_ = [("toneMapping", C.c_uint32), ("correctGamma", C.c_uint32), ("noDirect", C.c_uint32),
                ("noIndirect", C.c_uint32)]


class GIReservoirs(C.Structure):
    _fields_ = [("reservoir", C.c_void_p), ("reservoirPrev", C.c_void_p)]


class SceneArrays(C.Structure):
    _fields_ = [("materials", C.c_void_p), ("numMaterials", C.c_uint32),
                ("materialIndices", C.c_void_p), ("numMaterialIndices", C.c_uint32),
                ("vertices", C.c_void_p), ("numVertices", C.c_uint32),
                ("indices", C.c_void_p), ("numIndices", C.c_uint32),
                ("instances", C.c_void_p), ("numInstances", C.c_uint32),
                ("lights", C.c_void_p), ("numLights", C.c_uint32),
                ("lightTable", C.c_void_p), ("numLightTable", C.c_uint32),
                ("textures", C.c_void_p), ("numTextures", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("filter", C.c_uint32), ("pad", C.c_uint32),
                ("rgba8", C.c_void_p)]


class GrisSettings(C.Structure):
    _fields_ = [("shiftType", C.c_uint32), ("rrScale", C.c_float), ("temporalReuse", C.c_uint32),
                ("spatialReuse", C.c_uint32), ("cap", C.c_uint32), ("maxDepth", C.c_uint32),
                ("spatialNeighbors", C.c_uint32), ("spatialRadius", C.c_float), ("targetFunction", C.c_uint32),
                ("unbiasedReuse", C.c_uint32), ("guidedProposal", C.c_uint32)]


class MlpConfig(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("featuresPerLevel", C.c_uint32), ("log2TableSize", C.c_uint32),
                ("baseResolution", C.c_uint32), ("maxResolution", C.c_uint32), ("hiddenWidth", C.c_uint32),
                ("hiddenLayers", C.c_uint32), ("seed", C.c_uint32), ("learningRate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float), ("boundsLo", C.c_float * 3),
                ("boundsHi", C.c_float * 3)]


class Frame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("directOutput", C.c_void_p),
                ("indirectOutput", C.c_void_p), ("depthNormal", C.c_void_p), ("depthNormalPrev", C.c_void_p),
                ("albedoMatId", C.c_void_p), ("albedoMatIdPrev", C.c_void_p), ("motionVector", C.c_void_p),
                ("grisReservoir", C.c_void_p), ("grisReservoirPrev", C.c_void_p), ("grisReservoirTemp", C.c_void_p),
                ("guideLobes", C.c_void_p), ("guideLobesPrev", C.c_void_p), ("guideSamples", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("cams", Camera * 2), ("frame", Frame), ("settings", GrisSettings),
                ("target", C.c_void_p), ("guide", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("persistentGroups", C.c_uint32), ("xcdSwizzle", C.c_uint32), ("waves", C.c_uint32),
                ("queueTile", C.c_uint32), ("graph", C.c_uint32), ("shading", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("numPrimitives", C.c_uint32), ("numLights", C.c_uint32), ("bvhNodes", C.c_uint32),
                ("bvhDepth", C.c_uint32), ("sceneLo", C.c_float * 3), ("sceneHi", C.c_float * 3),
                ("deviceBytes", C.c_size_t), ("bvh4Nodes", C.c_uint32), ("bvh4Depth", C.c_uint32),
                ("bvhBuildMs", C.c_float), ("bvhBinaryMs", C.c_float)]


assert C.sizeof(Material) == 32 and C.sizeof(MeshVertex) == 32 and C.sizeof(ObjectInstance) == 224
assert C.sizeof(TriangleLight) == 64 and C.sizeof(Camera) == 352 and C.sizeof(GrisReservoir) == 96
assert C.sizeof(DIReservoir) == 64 and C.sizeof(DISettings) == 16


def settings(shift=2, rr_scale=1.0, temporal=False, spatial=True, cap=20, max_depth=15, neighbors=3, radius=20.0,
             learned=False, unbiased=False, guided=False):
    """GRISReSTIR::Settings (src/GRISReSTIR.h:29 defaults) + kernel constants."""
    return GrisSettings(shift, rr_scale, int(temporal), int(spatial), cap, max_depth, neighbors, radius, int(learned),
                        int(unbiased), int(guided))


def di_settings(shift=0, sample=0, temporal=False, spatial=True):
    """TestReSTIR::Settings; defaults are the reference's {Reconnection, Light, 0, 1} (src/TestReSTIR.h:31)."""
    return DISettings(shift, sample, int(temporal), int(spatial))


def mlp_config(lo, hi, log2_table=19, base_res=16, max_res=2048, seed=3, lr=1e-2):
    """Hash grid 16 x 2, 2^log2_table entries per level; MLP 2 x 64; Adam (0.9, 0.99, 1e-15)."""
    return MlpConfig(16, 2, log2_table, base_res, max_res, 64, 2, seed, lr, 0.9, 0.99, 1e-15,
                     (C.c_float * 3)(*lo), (C.c_float * 3)(*hi))


# ---------------------------------------------------------------------------------------
# Library loading

CORE_SYMBOLS = [
    "rt_last_error", "rt_device_scene_create", "rt_device_scene_destroy", "rt_device_scene_info", "rt_bvh_export", "rt_bvh4_export",
    "rt_bvh_build_noalloc",
    "rt_trace_rays", "rt_gbuffer_trace", "rt_pt_plain", "rt_gris_path_trace", "rt_gris_path_trace_groups",
    "rt_gris_temporal", "rt_gris_spatial",
    "rt_gris_render", "rt_gris_reservoir_bytes", "rt_gris_reservoir_reset", "rt_gris_reservoir_download",
    "rt_gris_reservoir_upload", "rt_gris_learned_download", "rt_gris_learned_upload", "rt_counters_enable",
    "rt_counters_read", "rt_mlp_create", "rt_mlp_destroy", "rt_mlp_param_counts", "rt_mlp_get_params",
    "rt_mlp_set_params", "rt_mlp_infer", "rt_mlp_gradients", "rt_mlp_read_gradients", "rt_mlp_adam_step",
    "rt_mlp_train_step", "rt_mlp_train_frame", "rt_mlp_batch_download", "rt_gris_learned_refresh",
    "rt_guide_create", "rt_guide_lobe_bytes", "rt_gris_guide_refresh", "rt_gris_guide_train",
    "rt_frame_graph_create", "rt_frame_graph_launch", "rt_frame_graph_destroy", "rt_batch_create",
    "rt_batch_destroy", "rt_batch_frame", "rt_batch_set_tuning", "rt_batch_get_tuning", "rt_batch_autotune",
    "rt_di_reservoir_bytes", "rt_di_reservoir_reset", "rt_di_reservoir_download", "rt_di_reservoir_upload",
    "rt_di_path_gen", "rt_di_temporal", "rt_di_spatial", "rt_di_render", "rt_post_process",
    "rt_gi_reservoir_bytes", "rt_gi_reservoir_reset", "rt_gi_reservoir_download", "rt_gi_reservoir_upload",
    "rt_gi_temporal",
]
HOST_SYMBOLS = [
    "rt_scenegen", "rt_scene_load", "rt_host_scene_film_size", "rt_host_scene_destroy", "rt_host_scene_arrays", "rt_host_scene_camera", "rt_light_table_build",
    "rt_renderer_create", "rt_renderer_destroy", "rt_renderer_set_settings", "rt_renderer_frame",
    "rt_renderer_bindings", "rt_renderer_flip", "rt_renderer_read_output", "rt_renderer_reset_reservoirs",
    "rt_renderer_enable_learned", "rt_renderer_enable_proposal", "rt_renderer_frame_methods",
    "rt_renderer_set_di_settings",
    "rt_renderer_di_bindings", "rt_write_png", "rt_read_png", "rt_renderer_screenshot",
    "rt_pipeline_create", "rt_pipeline_destroy", "rt_pipeline_set_settings", "rt_pipeline_set_trace_groups",
    "rt_pipeline_set_direct_method", "rt_pipeline_set_di_settings", "rt_pipeline_di_bindings",
    "rt_pipeline_set_direct_stream",
    "rt_pipeline_frame",
    "rt_pipeline_bindings", "rt_pipeline_read_output", "rt_pipeline_reset_reservoirs",
]

_core = None
_host = None


def _declare_core(L):
    P, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
    L.rt_last_error.restype = C.c_char_p
    sigs = {
        "rt_device_scene_create": [C.POINTER(SceneArrays), P, C.POINTER(P)],
        "rt_device_scene_destroy": [P],
        "rt_device_scene_info": [P, C.POINTER(SceneInfo)],
        "rt_bvh_export": [P, P, P, P],
        "rt_bvh4_export": [P, P],
        "rt_bvh_build_noalloc": [P, u32, P, P, P, P],
        "rt_trace_rays": [P, P, u32, C.c_int, C.c_int, P, P, P, P, P],
        "rt_gbuffer_trace": [P, C.POINTER(Camera), P, P, P, P],
        "rt_pt_plain": [P, C.POINTER(Camera), C.POINTER(Frame), u32, u32, P],
        "rt_gris_path_trace": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GrisSettings), P],
        "rt_gris_path_trace_groups": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GrisSettings), u32, P],
        "rt_gris_temporal": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GrisSettings), P],
        "rt_gris_spatial": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GrisSettings), P],
        "rt_gris_render": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GrisSettings), P],
        "rt_gris_reservoir_reset": [P, u32, u32, P],
        "rt_gris_reservoir_download": [P, u32, u32, P, P],
        "rt_gris_reservoir_upload": [P, u32, u32, P, P],
        "rt_gris_learned_download": [P, u32, u32, P, P],
        "rt_gris_learned_upload": [P, u32, u32, P, P],
        "rt_counters_enable": [C.c_int],
        "rt_counters_read": [P, C.c_int],
        "rt_mlp_create": [C.POINTER(MlpConfig), P, C.POINTER(P)],
        "rt_mlp_destroy": [P],
        "rt_mlp_param_counts": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "rt_mlp_get_params": [P, P, P],
        "rt_mlp_set_params": [P, P, P],
        "rt_mlp_infer": [P, P, u32, P, P],
        "rt_mlp_gradients": [P, P, P, u32, P, P],
        "rt_mlp_read_gradients": [P, P, P],
        "rt_mlp_adam_step": [P, P],
        "rt_mlp_train_step": [P, P, P, u32, P, P],
        "rt_mlp_train_frame": [P, P, P],
        "rt_mlp_batch_download": [P, P, P, C.POINTER(C.c_uint32)],
        "rt_gris_learned_refresh": [P, P, C.POINTER(Camera), C.POINTER(Frame), P],
        "rt_guide_create": [C.POINTER(MlpConfig), P, C.POINTER(P)],
        "rt_gris_guide_refresh": [P, P, C.POINTER(Camera), C.POINTER(Frame), P],
        "rt_gris_guide_train": [P, P, C.POINTER(Camera), C.POINTER(Frame), P, P],
        "rt_frame_graph_create": [P, C.POINTER(Frame), C.POINTER(GrisSettings), P, P, C.POINTER(P)],
        "rt_frame_graph_launch": [P, u32, C.POINTER(Camera), P],
        "rt_frame_graph_destroy": [P],
        "rt_batch_create": [u32, C.POINTER(P)],
        "rt_batch_destroy": [P],
        "rt_batch_frame": [P, C.POINTER(View), u32, P],
        "rt_batch_set_tuning": [P, C.POINTER(BatchTuning)],
        "rt_batch_get_tuning": [P, C.POINTER(BatchTuning)],
        "rt_batch_autotune": [P, C.POINTER(View), u32, u32, P, C.POINTER(BatchTuning), C.POINTER(C.c_float)],
        "rt_di_reservoir_reset": [P, u32, u32, P],
        "rt_di_reservoir_download": [P, u32, u32, P, P],
        "rt_di_reservoir_upload": [P, u32, u32, P, P],
        "rt_post_process": [C.POINTER(Frame), C.POINTER(PostSettings), P, P],
        "rt_gi_reservoir_reset": [P, u32, u32, P],
        "rt_gi_reservoir_download": [P, u32, u32, P, P],
        "rt_gi_reservoir_upload": [P, u32, u32, P, P],
        "rt_gi_temporal": [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(GIReservoirs), P],
    }
    di_pass = [P, C.POINTER(Camera), C.POINTER(Frame), C.POINTER(DIReservoirs), C.POINTER(DISettings), P]
    for name in ("rt_di_path_gen", "rt_di_temporal", "rt_di_spatial", "rt_di_render"):
        sigs[name] = di_pass
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32
    for name in ("rt_gris_reservoir_bytes", "rt_di_reservoir_bytes", "rt_gi_reservoir_bytes", "rt_guide_lobe_bytes"):
        getattr(L, name).argtypes = [u32, u32]
        getattr(L, name).restype = C.c_size_t


def _declare_host(L):
    P, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
    sigs = {
        "rt_scenegen": [u32, u32, C.POINTER(P)],
        "rt_scene_load": [C.c_char_p, C.POINTER(P)],
        "rt_host_scene_film_size": [P, C.POINTER(u32), C.POINTER(u32)],
        "rt_host_scene_destroy": [P],
        "rt_host_scene_arrays": [P, C.POINTER(SceneArrays)],
        "rt_host_scene_camera": [P, u32, u32, u32, C.POINTER(Camera)],
        "rt_light_table_build": [P, u32, P],
        "rt_pipeline_create": [P, u32, u32, C.POINTER(P)],
        "rt_pipeline_destroy": [P],
        "rt_pipeline_set_settings": [P, C.POINTER(GrisSettings)],
        "rt_pipeline_set_trace_groups": [P, u32],
        "rt_pipeline_frame": [P, C.POINTER(Camera), P, P],
        "rt_pipeline_bindings": [P, C.POINTER(Frame)],
        "rt_pipeline_set_direct_method": [P, u32],
        "rt_pipeline_set_di_settings": [P, C.POINTER(DISettings)],
        "rt_pipeline_di_bindings": [P, C.POINTER(DIReservoirs)],
        "rt_pipeline_set_direct_stream": [P, P],
        "rt_pipeline_read_output": [P, u32, P, P],
        "rt_pipeline_reset_reservoirs": [P, P],
        "rt_renderer_create": [P, u32, u32, C.POINTER(P)],
        "rt_renderer_destroy": [P],
        "rt_renderer_set_settings": [P, C.POINTER(GrisSettings)],
        "rt_renderer_frame": [P, C.POINTER(Camera), u32, P],
        "rt_renderer_bindings": [P, C.POINTER(Frame)],
        "rt_renderer_flip": [P],
        "rt_renderer_read_output": [P, u32, P, P],
        "rt_renderer_reset_reservoirs": [P, P],
        "rt_renderer_enable_learned": [P, u32, u32, P, C.POINTER(P)],
        "rt_renderer_enable_proposal": [P, u32, u32, P, C.POINTER(P)],
        "rt_renderer_frame_methods": [P, C.POINTER(Camera), u32, u32, P],
        "rt_renderer_set_di_settings": [P, C.POINTER(DISettings)],
        "rt_renderer_di_bindings": [P, C.POINTER(DIReservoirs)],
        "rt_write_png": [C.c_char_p, P, u32, u32],
        "rt_read_png": [C.c_char_p, C.POINTER(u32), C.POINTER(u32), P],
        "rt_renderer_screenshot": [P, C.POINTER(PostSettings), C.c_char_p, P],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32


def core():
    """librestirpt.so (HIP kernels + C ABI). Raises if it has not been built."""
    global _core
    if _core is None:
        if not os.path.exists(CORE_LIB):
            raise RuntimeError(f"{CORE_LIB} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(CORE_LIB, mode=C.RTLD_GLOBAL)
        _declare_core(L)
        _core = L
    return _core


def host():
    """librestirpt_host.so (C++ pass surface, scene generator, frame driver)."""
    global _host
    if _host is None:
        core()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing: build it first")
        L = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
        _declare_host(L)
        _host = L
    return _host


def check(status, what="rt call"):
    if status != RT_OK:
        msg = core().rt_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with status {status}: {msg}")


# ---------------------------------------------------------------------------------------
# Host scene (procedural workloads) with numpy views of its arrays


def _view(ptr, count, ctype):
    if count == 0:
        return np.zeros(0, dtype=np.dtype(ctype))
    arr = (ctype * count).from_address(ptr)
    return np.ctypeslib.as_array(arr)


class HostScene:
    """Host scene: procedural (rt_scenegen: config 0 box room, 1 terrain, 2 floor + light, ...) or
    loaded from a scene file in the reference's XML format (HostScene.load, Scene::load)."""

    def __init__(self, config: int, seed: int = 0, _handle=None):
        H = host()
        h = C.c_void_p()
        if _handle is None:
            check(H.rt_scenegen(config, seed, C.byref(h)), "rt_scenegen")
        else:
            h = _handle
        self.handle = h
        self.arrays = SceneArrays()
        check(H.rt_host_scene_arrays(h, C.byref(self.arrays)), "rt_host_scene_arrays")

    @classmethod
    def load(cls, path):
        """rt_scene_load: the scene file at `path` (XML + OBJ meshes beside it)."""
        h = C.c_void_p()
        check(host().rt_scene_load(str(path).encode(), C.byref(h)), "rt_scene_load")
        return cls(0, 0, _handle=h)

    def film_size(self):
        w, h = C.c_uint32(), C.c_uint32()
        check(host().rt_host_scene_film_size(self.handle, C.byref(w), C.byref(h)), "rt_host_scene_film_size")
        return w.value, h.value

    def camera(self, frame: int, width: int, height: int):
        cams = (Camera * 2)()
        check(host().rt_host_scene_camera(self.handle, frame, width, height, cams), "rt_host_scene_camera")
        return cams

    def numpy(self):
        a = self.arrays
        return {
            "materials": _view(a.materials, a.numMaterials, Material),
            "materialIndices": _view(a.materialIndices, a.numMaterialIndices, C.c_int32),
            "vertices": _view(a.vertices, a.numVertices, MeshVertex),
            "indices": _view(a.indices, a.numIndices, C.c_uint32),
            "instances": _view(a.instances, a.numInstances, ObjectInstance),
            "lights": _view(a.lights, a.numLights, TriangleLight),
            "lightTable": _view(a.lightTable, a.numLightTable, LightTableEntry),
        }

    def num_triangles(self):
        n = self.arrays.numLights
        inst = _view(self.arrays.instances, self.arrays.numInstances, ObjectInstance)
        return n + int(sum(int(i["indexCount"]) // 3 for i in inst))

    def __del__(self):
        if getattr(self, "handle", None) and _host is not None:
            _host.rt_host_scene_destroy(self.handle)
            self.handle = None


class DeviceScene:
    """rt_device_scene: buffers uploaded and LBVH built on the current device."""

    def __init__(self, arrays: SceneArrays, stream=None):
        L = core()
        h = C.c_void_p()
        check(L.rt_device_scene_create(C.byref(arrays), stream, C.byref(h)), "rt_device_scene_create")
        self.handle = h

    def info(self) -> SceneInfo:
        i = SceneInfo()
        check(core().rt_device_scene_info(self.handle, C.byref(i)), "rt_device_scene_info")
        return i

    def __del__(self):
        if getattr(self, "handle", None) and _core is not None:
            _core.rt_device_scene_destroy(self.handle)
            self.handle = None


class Mlp:
    """rt_mlp: the learned target function (hash grid + MLP) on the current device; with
    guide=True the learned proposal's network (rt_guide_create: 32 outputs, an 8-lobe vMF mixture)."""

    def __init__(self, config: MlpConfig, stream=None, guide=False):
        L = core()
        h = C.c_void_p()
        if guide:
            check(L.rt_guide_create(C.byref(config), stream, C.byref(h)), "rt_guide_create")
        else:
            check(L.rt_mlp_create(C.byref(config), stream, C.byref(h)), "rt_mlp_create")
        self.guide = guide
        self.handle = h
        self.config = config
        g, w = C.c_uint64(), C.c_uint64()
        check(L.rt_mlp_param_counts(h, C.byref(g), C.byref(w)), "rt_mlp_param_counts")
        self.grid_params, self.weight_params = g.value, w.value

    def params(self):
        grid = np.zeros(self.grid_params, np.float32)
        w = np.zeros(self.weight_params, np.float32)
        check(core().rt_mlp_get_params(self.handle, grid.ctypes.data, w.ctypes.data), "rt_mlp_get_params")
        return grid, w

    def set_params(self, grid=None, weights=None):
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        check(core().rt_mlp_set_params(self.handle, ptr(g), ptr(w)), "rt_mlp_set_params")

    def read_gradients(self):
        grid = np.zeros(self.grid_params, np.float32)
        w = np.zeros(self.weight_params, np.float32)
        check(core().rt_mlp_read_gradients(self.handle, grid.ctypes.data, w.ctypes.data), "rt_mlp_read_gradients")
        return grid, w

    def batch(self):
        """Training batch gathered by the last rt_gris_learned_refresh (rt_gris_guide_train):
        (pos [n,3], target [n,3] or [n,4])."""
        pos = np.zeros((1 << 16, 3), np.float32)
        tgt = np.zeros((1 << 16, 4 if self.guide else 3), np.float32)
        n = C.c_uint32()
        check(core().rt_mlp_batch_download(self.handle, pos.ctypes.data, tgt.ctypes.data, C.byref(n)), "batch")
        return pos[: n.value], tgt[: n.value]

    def __del__(self):
        if getattr(self, "handle", None) and _core is not None:
            _core.rt_mlp_destroy(self.handle)
            self.handle = None


def ptr(t):
    """Device (or host) pointer of a torch tensor / numpy array."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()
