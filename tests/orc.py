"""ctypes binding of the oracle (oracle/liboracle.so) for the tests.

The oracle is the scalar C++ restatement of the reference GLSL; it is only
ever used as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import rtpkg

rt = rtpkg.rt
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = C.CDLL(ORACLE_LIB)
        P, u32 = C.c_void_p, C.c_uint32
        L.orc_scene_create.argtypes = [C.POINTER(rt.SceneArrays)]
        L.orc_scene_create.restype = P
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_num_prims.argtypes = [P]
        L.orc_scene_num_prims.restype = u32
        L.orc_scene_world_tris.argtypes = [P, P, P]
        L.orc_bvh_export.argtypes = [P, P, P, P, P]
        L.orc_bvh_num_nodes.argtypes = [P]
        L.orc_bvh_num_nodes.restype = u32
        L.orc_bvh4_num_nodes.argtypes = [P]
        L.orc_bvh4_num_nodes.restype = u32
        L.orc_bvh4_export.argtypes = [P, P]
        L.orc_sample_texture.argtypes = [C.POINTER(rt.Texture), P, u32, P]
        L.orc_srgb_table.argtypes = [P]
        L.orc_trace_rays.argtypes = [P, P, u32, C.c_int, C.c_int, P, P, P, P, P, C.c_int]
        L.orc_gbuffer.argtypes = [P, C.POINTER(rt.Camera), P, P, P, P, C.c_int]
        L.orc_pt_plain.argtypes = [P, C.POINTER(rt.Camera), P, P, u32, P, P, P, C.c_int]
        L.orc_gris_path_trace.argtypes = [P, C.POINTER(rt.Camera), P, P, C.POINTER(rt.GrisSettings), P, P, C.c_int]
        L.orc_gris_temporal.argtypes = [P, C.POINTER(rt.Camera), P, P, P, P, P, C.POINTER(rt.GrisSettings), P, P, P,
                                        P, C.c_int, P, P, P]
        L.orc_gris_spatial.argtypes = [P, C.POINTER(rt.Camera), P, P, C.POINTER(rt.GrisSettings), P, P, P, P, C.c_int,
                                       P, P]
        L.orc_di_path_gen.argtypes = [P, C.POINTER(rt.Camera), P, P, C.POINTER(rt.DISettings), P, P, C.c_int]
        L.orc_di_temporal.argtypes = [P, C.POINTER(rt.Camera), P, P, P, P, P, C.POINTER(rt.DISettings), P, P, P, P,
                                      C.c_int]
        L.orc_di_spatial.argtypes = [P, C.POINTER(rt.Camera), P, P, C.POINTER(rt.DISettings), P, P, P, P, C.c_int]
        L.orc_gi_temporal.argtypes = [P, C.POINTER(rt.Camera), P, P, P, P, P, P, P, P, P, C.c_int]
        L.orc_post_process.argtypes = [P, P, u32, u32, P, P]
        L.orc_tonemap.argtypes = [C.c_float, u32]
        L.orc_tonemap.restype = C.c_float
        L.orc_gris_learned_refresh.argtypes = [P, u32, u32, P, P, P, P, P, P, u32, u32, C.c_int]
        L.orc_mlp_forward.argtypes = [u32, u32, u32, P, P, P, P, P, u32, P, C.c_int]
        L.orc_mlp_gradients.argtypes = [u32, u32, u32, P, P, P, P, P, P, u32, P, P, C.c_int]
        L.orc_mlp_gradients.restype = C.c_double
        L.orc_mlp_loss.argtypes = [u32, u32, u32, P, P, P, P, P, P, u32, P, P]
        L.orc_mlp_loss.restype = C.c_double
        L.orc_mlp_levels.argtypes = [u32, u32, u32, P]
        L.orc_gris_bind_guide.argtypes = [P, P, P, P]
        L.orc_guide_mixture.argtypes = [P, P, u32, P, P, u32, P]
        L.orc_explog.argtypes = [P, u32, C.c_int, P]
        L.orc_guide_forward.argtypes = [u32, u32, u32, P, P, P, P, P, u32, P, C.c_int]
        L.orc_guide_lobes.argtypes = [P, u32, P]
        L.orc_guide_gradients.argtypes = [u32, u32, u32, P, P, P, P, P, P, u32, P, P, C.c_int]
        L.orc_guide_gradients.restype = C.c_double
        L.orc_guide_loss.argtypes = [u32, u32, u32, P, P, P, P, P, P, u32, C.c_double]
        L.orc_guide_loss.restype = C.c_double
        L.orc_mlp_levels.restype = u32
        L.orc_hash2.argtypes = [u32]
        L.orc_hash2.restype = u32
        L.orc_make_seed.argtypes = [u32, u32, u32]
        L.orc_make_seed.restype = u32
        L.orc_sample1f.argtypes = [u32, u32, P]
        L.orc_sincos.argtypes = [P, u32, P, P]
        L.orc_concentric.argtypes = [P, u32, P]
        L.orc_camera_rays.argtypes = [C.POINTER(rt.Camera), P]
        L.orc_eval_bsdf.argtypes = [C.POINTER(rt.Material), P, P, u32, P, P]
        L.orc_sample_bsdf.argtypes = [C.POINTER(rt.Material), P, P, u32, P, P, P, P, P]
        L.orc_gtr2_sample.argtypes = [P, u32, C.c_float, P]
        L.orc_sample_light.argtypes = [C.POINTER(rt.SceneArrays), P, u32, P, P]
        L.orc_alias_build.argtypes = [P, u32, P]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        _lib = L
    return _lib


def nthreads():
    return max(1, min(8, os.cpu_count() or 1))


def p(a):
    return None if a is None else a.ctypes.data


class OracleScene:
    def __init__(self, arrays):
        self.h = lib().orc_scene_create(C.byref(arrays))
        self.n = lib().orc_scene_num_prims(self.h)

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h and _lib is not None:
            try:
                _lib.orc_scene_destroy(h)
            except Exception:
                pass

    def bvh(self):
        n = self.n
        nn = lib().orc_bvh_num_nodes(self.h)
        codes = np.zeros(n, np.uint64)
        order = np.zeros(n, np.uint32)
        nodes = np.zeros((nn, 14), np.uint32)
        box = np.zeros(6, np.float32)
        lib().orc_bvh_export(self.h, p(codes), p(order), p(nodes), p(box))
        return codes, order, nodes, box

    def bvh4(self):
        nn = lib().orc_bvh4_num_nodes(self.h)
        nodes = np.zeros((nn, 28), np.uint32)
        lib().orc_bvh4_export(self.h, p(nodes))
        return nodes

    def world_tris(self):
        pos = np.zeros((self.n, 9), np.float32)
        ref = np.zeros((self.n, 2), np.uint32)
        lib().orc_scene_world_tris(self.h, p(pos), p(ref))
        return pos, ref

    def trace(self, rays, mode=0, skip_lights=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        bary = np.zeros((n, 2), np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        stats = np.zeros(3, np.uint64)
        lib().orc_trace_rays(self.h, p(rays), n, mode, int(skip_lights), p(t), p(bary), p(inst), p(prim), p(stats),
                             nthreads())
        return t, bary, inst, prim, stats

    def gbuffer(self, cam):
        W, H = cam.filmSize[0], cam.filmSize[1]
        dn = np.zeros((H, W, 4), np.float32)
        am = np.zeros((H, W, 2), np.uint32)
        mv = np.zeros((H, W, 2), np.uint16)
        st = np.zeros(3, np.uint64)
        lib().orc_gbuffer(self.h, C.byref(cam), p(dn), p(am), p(mv), p(st), nthreads())
        return dn, am, mv, st

    def pt_plain(self, cam, dn, am, max_depth, direct=None, indirect=None):
        W, H = cam.filmSize[0], cam.filmSize[1]
        direct = np.zeros((H, W, 4), np.float32) if direct is None else direct
        indirect = np.zeros((H, W, 4), np.float32) if indirect is None else indirect
        st = np.zeros(3, np.uint64)
        lib().orc_pt_plain(self.h, C.byref(cam), p(dn), p(am), max_depth, p(direct), p(indirect), p(st), nthreads())
        return direct, indirect, st

    def gris_path_trace(self, cams, dn, am, settings, resv):
        st = np.zeros(3, np.uint64)
        lib().orc_gris_path_trace(self.h, cams, p(dn), p(am), C.byref(settings), p(resv), p(st), nthreads())
        return st

    def gris_temporal(self, cams, dn, am, mv, dn_prev, am_prev, settings, resv_this, resv_prev, resv_temp,
                      learn_this=None, learn_prev=None, learn_temp=None):
        st = np.zeros(3, np.uint64)
        lib().orc_gris_temporal(self.h, cams, p(dn), p(am), p(mv), p(dn_prev), p(am_prev), C.byref(settings),
                                p(resv_this), p(resv_prev), p(resv_temp), p(st), nthreads(), p(learn_this),
                                p(learn_prev), p(learn_temp))
        return st

    def gris_spatial(self, cams, dn, am, settings, resv_temp, resv_this, indirect, learn_temp=None, learn_this=None):
        st = np.zeros(3, np.uint64)
        lib().orc_gris_spatial(self.h, cams, p(dn), p(am), C.byref(settings), p(resv_temp), p(resv_this),
                               p(indirect), p(st), nthreads(), p(learn_temp), p(learn_this))
        return st

    def di_path_gen(self, cams, dn, am, settings, resv):
        st = np.zeros(3, np.uint64)
        lib().orc_di_path_gen(self.h, cams, p(dn), p(am), C.byref(settings), p(resv), p(st), nthreads())
        return st

    def di_temporal(self, cams, dn, am, mv, dn_prev, am_prev, settings, resv_this, resv_prev, resv_temp):
        st = np.zeros(3, np.uint64)
        lib().orc_di_temporal(self.h, cams, p(dn), p(am), p(mv), p(dn_prev), p(am_prev), C.byref(settings),
                              p(resv_this), p(resv_prev), p(resv_temp), p(st), nthreads())
        return st

    def di_spatial(self, cams, dn, am, settings, resv_temp, resv_this, direct):
        st = np.zeros(3, np.uint64)
        lib().orc_di_spatial(self.h, cams, p(dn), p(am), C.byref(settings), p(resv_temp), p(resv_this), p(direct),
                             p(st), nthreads())
        return st

    def gi_temporal(self, cams, dn, am, mv, dn_prev, am_prev, resv_this, resv_prev, indirect):
        st = np.zeros(3, np.uint64)
        lib().orc_gi_temporal(self.h, cams, p(dn), p(am), p(mv), p(dn_prev), p(am_prev), p(resv_this), p(resv_prev),
                              p(indirect), p(st), nthreads())
        return st

    def bind_guide(self, lobes=None, lobes_prev=None, samples=None):
        """Learned-proposal planes for the following GRIS calls (arrays kept alive by the caller):
        lobes / lobes_prev float32 [16, H*W, 4], samples float32 [H*W, 4] (written by the path tracer)."""
        self._guide = (lobes, lobes_prev, samples)
        lib().orc_gris_bind_guide(self.h, p(lobes), p(lobes_prev), p(samples))

    def learned_refresh(self, W, H, dn, resv, lhat4, stride, offset):
        """Refreshes resv in place from the predictions lhat4 (H*W x 4); returns (learned planes, batch pos,
        batch target) with the batch over pixels p % stride == offset."""
        n = W * H
        learn = np.zeros((n, 4), np.float32)
        slots = (n - offset + stride - 1) // stride
        bpos = np.zeros((slots, 3), np.float32)
        btgt = np.zeros((slots, 3), np.float32)
        lib().orc_gris_learned_refresh(self.h, W, H, p(dn), p(resv), p(np.ascontiguousarray(lhat4, np.float32)),
                                       p(learn), p(bpos), p(btgt), stride, offset, nthreads())
        return learn, bpos, btgt


RESV_DTYPE = np.dtype([("bary", np.float32, 2), ("instanceIdx", np.uint32), ("triangleIdx", np.uint32),
                       ("rcLi", np.float32, 3), ("rcRng", np.uint32), ("rcWi", np.float32, 3), ("flags", np.uint32),
                       ("pad", np.float32, 2), ("rcPrevSamplePdf", np.float32), ("rcJacobian", np.float32),
                       ("F", np.float32, 3), ("primaryRng", np.uint32), ("sampleCount", np.float32),
                       ("resampleWeight", np.float32), ("contribWeight", np.float32), ("pad0", np.float32)])
assert RESV_DTYPE.itemsize == 96


def post_process(direct, indirect, settings4):
    """direct/indirect: (H, W, 4) float32; settings4: (toneMapping, correctGamma, noDirect, noIndirect)."""
    H, W = direct.shape[:2]
    out = np.zeros((H, W, 4), np.uint8)
    st = np.asarray(settings4, np.uint32)
    lib().orc_post_process(p(np.ascontiguousarray(direct, np.float32)), p(np.ascontiguousarray(indirect, np.float32)),
                           W, H, p(st), p(out))
    return out


DI_DTYPE = np.dtype([("bary", np.float32, 2), ("instanceIdx", np.uint32), ("triangleIdx", np.uint32),
                     ("Li", np.float32, 3), ("pad0", np.float32), ("jacobian", np.float32), ("samplePdf", np.float32),
                     ("rng", np.uint32), ("isLightSample", np.uint32), ("sampleCount", np.uint32),
                     ("resampleWeight", np.float32), ("contribWeight", np.float32), ("weight", np.float32)])
assert DI_DTYPE.itemsize == 64


GI_DTYPE = np.dtype([("bary", np.float32, 2), ("instanceIdx", np.uint32), ("triangleIdx", np.uint32),
                     ("rcLo", np.float32, 3), ("rcPrevCoord", np.uint32), ("sampleCount", np.uint32),
                     ("resampleWeight", np.float32), ("contribWeight", np.float32), ("pad0", np.float32)])
assert GI_DTYPE.itemsize == 48


def reset_gi_reservoirs(n):
    """The state rt_gi_reservoir_reset writes: zeros with an invalid sample."""
    r = np.zeros(n, GI_DTYPE)
    r["instanceIdx"] = 0xFFFFFFFF
    return r


def reset_di_reservoirs(n):
    """The state rt_di_reservoir_reset writes: zeros with an invalid sample."""
    r = np.zeros(n, DI_DTYPE)
    r["instanceIdx"] = 0xFFFFFFFF
    return r


class OracleMlp:
    """Scalar restatement of the hash-grid MLP (mode 1: FP16 rounding as the kernels; mode 0: exact)."""

    def __init__(self, log2_table, base_res, max_res, lo, hi):
        self.log2, self.base, self.max = log2_table, base_res, max_res
        self.lo = np.asarray(lo, np.float32)
        self.hi = np.asarray(hi, np.float32)

    def levels(self):
        res = np.zeros(16, np.uint32)
        dense = lib().orc_mlp_levels(self.log2, self.base, self.max, p(res))
        return res, dense

    def forward(self, grid, weights, pos, mode=1):
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros((len(pos), 3), np.float32)
        lib().orc_mlp_forward(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights), p(pos),
                              len(pos), p(out), mode)
        return out

    def gradients(self, grid, weights, pos, target, mode=1):
        pos = np.ascontiguousarray(pos, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        gg = np.zeros(len(grid), np.float64)
        gw = np.zeros(len(weights), np.float64)
        loss = lib().orc_mlp_gradients(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights),
                                       p(pos), p(target), len(pos), p(gg), p(gw), mode)
        return loss, gg, gw

    def loss(self, grid, weights, pos, target, den_in=None, den_out=None):
        pos = np.ascontiguousarray(pos, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        return lib().orc_mlp_loss(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights), p(pos),
                                  p(target), len(pos), p(den_in), p(den_out))


class OracleGuide(OracleMlp):
    """The learned proposal's network: 32 raw outputs per position, the vMF-mixture loss."""

    def forward(self, grid, weights, pos, mode=1):
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.zeros((len(pos), 32), np.float32)
        lib().orc_guide_forward(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights), p(pos),
                                len(pos), p(out), mode)
        return out

    def gradients(self, grid, weights, pos, target, mode=1):
        pos = np.ascontiguousarray(pos, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        gg = np.zeros(len(grid), np.float64)
        gw = np.zeros(len(weights), np.float64)
        loss = lib().orc_guide_gradients(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights),
                                         p(pos), p(target), len(pos), p(gg), p(gw), mode)
        return loss, gg, gw

    def loss(self, grid, weights, pos, target, wmean=0.0):
        pos = np.ascontiguousarray(pos, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        return lib().orc_guide_loss(self.log2, self.base, self.max, p(self.lo), p(self.hi), p(grid), p(weights),
                                    p(pos), p(target), len(pos), wmean)


def guide_lobes(raw):
    """Lobe planes of raw proposal outputs: [n, 16, 4] (per sample the kernels' 16 float4)."""
    raw = np.ascontiguousarray(raw, np.float32)
    out = np.zeros((len(raw), 16, 4), np.float32)
    lib().orc_guide_lobes(p(raw), len(raw), p(out))
    return out


def mixture(lobes16x4, dirs, u3):
    """One pixel's mixture: pdf at dirs [n, 3] and samples for u3 [m, 3] (lobe choice, u1, u2)."""
    lobes = np.ascontiguousarray(lobes16x4, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    u3 = np.ascontiguousarray(u3, np.float32).reshape(-1, 3)
    pdf = np.zeros(len(dirs), np.float32)
    out = np.zeros((len(u3), 3), np.float32)
    lib().orc_guide_mixture(p(lobes), p(dirs), len(dirs), p(pdf), p(u3), len(u3), p(out))
    return pdf, out


def explog(x, log=False):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    lib().orc_explog(p(x), x.size, int(log), p(y))
    return y


def reset_reservoirs(n):
    r = np.zeros(n, RESV_DTYPE)
    r["instanceIdx"] = 0xFFFFFFFF
    return r


RESV_FIELDS = ["bary", "instanceIdx", "triangleIdx", "rcLi", "rcRng", "rcWi", "flags", "rcPrevSamplePdf",
               "rcJacobian", "F", "primaryRng", "sampleCount", "resampleWeight"]
