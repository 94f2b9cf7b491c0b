"""GPU-side helpers for the tests: torch-allocated frame buffers driven through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import orc
import rtpkg

rt = rtpkg.rt


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def dev_empty(nbytes):
    return torch.zeros(max(1, nbytes), dtype=torch.uint8, device="cuda")


def to_host(t, dtype, shape):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype).reshape(shape).copy()


def from_host(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


class GrisRun:
    """Frame buffers of one view, bound like Renderer::initDescriptor (two
    G-buffer parities, two reservoir parities, one Temp, outputs)."""

    def __init__(self, ds, W, H):
        self.ds, self.W, self.H = ds, W, H
        n = W * H
        self.dn = [dev_empty(16 * n), dev_empty(16 * n)]
        self.am = [dev_empty(8 * n), dev_empty(8 * n)]
        self.mv = dev_empty(4 * n)
        rb = rt.core().rt_gris_reservoir_bytes(W, H)
        self.resv = [dev_empty(rb), dev_empty(rb)]
        self.temp = dev_empty(rb)
        self.direct = dev_empty(16 * n)
        self.indirect = dev_empty(16 * n)
        for b in self.resv + [self.temp]:
            rt.check(rt.core().rt_gris_reservoir_reset(b.data_ptr(), W, H, stream()), "reset")
        self.cur = 0

    def frame_struct(self):
        f = rt.Frame()
        f.width, f.height = self.W, self.H
        c, l = self.cur, self.cur ^ 1
        f.directOutput, f.indirectOutput = self.direct.data_ptr(), self.indirect.data_ptr()
        f.depthNormal, f.depthNormalPrev = self.dn[c].data_ptr(), self.dn[l].data_ptr()
        f.albedoMatId, f.albedoMatIdPrev = self.am[c].data_ptr(), self.am[l].data_ptr()
        f.motionVector = self.mv.data_ptr()
        f.grisReservoir, f.grisReservoirPrev = self.resv[c].data_ptr(), self.resv[l].data_ptr()
        f.grisReservoirTemp = self.temp.data_ptr()
        return f

    def gbuffer(self, cam):
        rt.check(rt.core().rt_gbuffer_trace(self.ds.handle, C.byref(cam), self.dn[self.cur].data_ptr(),
                                            self.am[self.cur].data_ptr(), self.mv.data_ptr(), stream()), "gbuffer")

    def gbuffer_host(self, parity=None):
        c = self.cur if parity is None else parity
        W, H = self.W, self.H
        return (to_host(self.dn[c], np.float32, (H, W, 4)), to_host(self.am[c], np.uint32, (H, W, 2)),
                to_host(self.mv, np.uint16, (H, W, 2)))

    def reservoirs(self, buf):
        out = np.zeros(self.W * self.H, orc.RESV_DTYPE)
        torch.cuda.synchronize()
        rt.check(rt.core().rt_gris_reservoir_download(buf.data_ptr(), self.W, self.H, out.ctypes.data, stream()),
                 "download")
        return out

    def path_trace(self, cams, s):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_path_trace(self.ds.handle, cams, C.byref(f), C.byref(s), stream()), "path trace")

    def temporal(self, cams, s):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_temporal(self.ds.handle, cams, C.byref(f), C.byref(s), stream()), "temporal")

    def spatial(self, cams, s):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_spatial(self.ds.handle, cams, C.byref(f), C.byref(s), stream()), "spatial")

    def plain(self, cams, max_depth, which=3):
        f = self.frame_struct()
        rt.check(rt.core().rt_pt_plain(self.ds.handle, cams, C.byref(f), max_depth, which, stream()), "pt plain")

    def learned_refresh(self, mlp, cams):
        f = self.frame_struct()
        rt.check(rt.core().rt_gris_learned_refresh(self.ds.handle, mlp.handle, cams, C.byref(f), stream()), "refresh")

    def learned(self, buf):
        """Learned-target plane of a reservoir buffer: (W*H, 4) float32."""
        out = np.zeros((self.W * self.H, 4), np.float32)
        torch.cuda.synchronize()
        rt.check(rt.core().rt_gris_learned_download(buf.data_ptr(), self.W, self.H, out.ctypes.data, stream()),
                 "learned download")
        return out

    def output(self, which=1):
        t = self.indirect if which == 1 else self.direct
        return to_host(t, np.float32, (self.H, self.W, 4))

    def frame(self, cams, s):
        """G-buffer + the three GRIS passes of one frame; returns the indirect output (no flip)."""
        self.gbuffer(cams[0])
        self.path_trace(cams, s)
        self.temporal(cams, s)
        self.spatial(cams, s)
        return self.output(1)

    def flip(self):
        self.cur ^= 1


def random_rays(n, lo, hi, seed=0, tmin=1e-4, tmax=1e7):
    g = np.random.default_rng(seed)
    o = g.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d.astype(np.float32)
    rays[:, 6] = tmin
    rays[:, 7] = tmax
    return rays


def gpu_trace(ds, rays, mode=0, skip_lights=False):
    n = rays.shape[0]
    d_rays = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    t = torch.zeros(n, dtype=torch.float32, device="cuda")
    bary = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    inst = torch.zeros(n, dtype=torch.int32, device="cuda")
    prim = torch.zeros(n, dtype=torch.int32, device="cuda")
    rt.check(rt.core().rt_trace_rays(ds.handle, d_rays.data_ptr(), n, mode, int(skip_lights), t.data_ptr(),
                                     bary.data_ptr(), inst.data_ptr(), prim.data_ptr(), stream()), "trace rays")
    torch.cuda.synchronize()
    return (t.cpu().numpy(), bary.cpu().numpy().reshape(n, 2), inst.cpu().numpy().view(np.uint32),
            prim.cpu().numpy().view(np.uint32))


def gpu_bvh(ds):
    info = ds.info()
    n = info.numPrimitives
    codes = np.zeros(n, np.uint64)
    order = np.zeros(n, np.uint32)
    nodes = np.zeros((max(1, n - 1), 14), np.uint32)
    rt.check(rt.core().rt_bvh_export(ds.handle, codes.ctypes.data, order.ctypes.data, nodes.ctypes.data), "export")
    return codes, order, nodes, info


def gpu_bvh4(ds):
    info = ds.info()
    nodes = np.zeros((max(1, info.bvh4Nodes), 28), np.uint32)
    rt.check(rt.core().rt_bvh4_export(ds.handle, nodes.ctypes.data), "export4")
    return nodes[: info.bvh4Nodes]


def compare_reservoirs(a, b):
    """Mask of pixels whose reservoir fields differ (floats compared by value, NaN == NaN)."""
    bad = np.zeros(a.shape[0], bool)
    for f in orc.RESV_FIELDS:
        x = a[f].reshape(a.shape[0], -1)
        y = b[f].reshape(a.shape[0], -1)
        if x.dtype == np.float32:
            same = (x == y) | (np.isnan(x) & np.isnan(y))
        else:
            same = x == y
        bad |= ~same.all(axis=1)
    return bad


def same_float(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b))


def still_cams(base, i, seed):
    """Camera pair for frame i of a still view: base[0] with a per-frame seed, frameIndex i (clear
    flag on frame 0), previous camera = itself (no motion)."""
    c = (rt.Camera * 2)()
    C.memmove(C.addressof(c), C.addressof(base), C.sizeof(rt.Camera))
    h = (seed * 0x9E3779B1 + i * 0x85EBCA77 + 0x165667B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & 0xFFFFFFFF
    c[0].seed = h ^ (h >> 12)
    c[0].frameIndex = 0x80000000 if i == 0 else i
    for k in range(16):
        c[0].lastProjView[k] = c[0].projView[k]
    C.memmove(C.addressof(c) + C.sizeof(rt.Camera), C.addressof(c), C.sizeof(rt.Camera))
    return c


class DIRun(GrisRun):
    """GrisRun plus the three DI reservoir buffers of descriptor bindings 7/8/9 (two parities + Temp)."""

    def __init__(self, ds, W, H):
        super().__init__(ds, W, H)
        rb = rt.core().rt_di_reservoir_bytes(W, H)
        self.di = [dev_empty(rb), dev_empty(rb)]
        self.di_temp = dev_empty(rb)
        for b in self.di + [self.di_temp]:
            rt.check(rt.core().rt_di_reservoir_reset(b.data_ptr(), W, H, stream()), "di reset")

    def di_struct(self):
        r = rt.DIReservoirs()
        r.reservoir, r.reservoirPrev = self.di[self.cur].data_ptr(), self.di[self.cur ^ 1].data_ptr()
        r.reservoirTemp = self.di_temp.data_ptr()
        return r

    def _di(self, name, cams, s):
        f = self.frame_struct()
        r = self.di_struct()
        rt.check(getattr(rt.core(), name)(self.ds.handle, cams, C.byref(f), C.byref(r), C.byref(s), stream()), name)

    def di_path_gen(self, cams, s):
        self._di("rt_di_path_gen", cams, s)

    def di_temporal(self, cams, s):
        self._di("rt_di_temporal", cams, s)

    def di_spatial(self, cams, s):
        self._di("rt_di_spatial", cams, s)

    def di_render(self, cams, s):
        self._di("rt_di_render", cams, s)

    def di_reservoirs(self, buf):
        out = np.zeros(self.W * self.H, orc.DI_DTYPE)
        torch.cuda.synchronize()
        rt.check(rt.core().rt_di_reservoir_download(buf.data_ptr(), self.W, self.H, out.ctypes.data, stream()),
                 "di download")
        return out


DI_FLOAT_FIELDS = ["bary", "Li", "pad0", "jacobian", "samplePdf", "resampleWeight", "contribWeight", "weight"]
DI_UINT_FIELDS = ["instanceIdx", "triangleIdx", "rng", "isLightSample", "sampleCount"]


def compare_di(a, b):
    """Mask of pixels whose DI reservoirs differ (floats by value with NaN == NaN, integers exactly)."""
    bad = np.zeros(a.shape[0], bool)
    for f in DI_FLOAT_FIELDS:
        x, y = a[f].reshape(a.shape[0], -1), b[f].reshape(a.shape[0], -1)
        bad |= ~((x == y) | (np.isnan(x) & np.isnan(y))).all(axis=1)
    for f in DI_UINT_FIELDS:
        bad |= a[f] != b[f]
    return bad
