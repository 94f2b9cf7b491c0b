"""Benchmark: ReSTIR PT (GRIS) frames at 1080p on one MI355X.

A step is one frame of the hot path over one 1920x1080 view: G-buffer
(primary rays), GRIS initial path generation, temporal reuse and spatial reuse
(hybrid shift, 3 neighbours in 30 px, M-cap 20, depth 6, 1 path/px) on the
procedural ~1.11M-triangle terrain scene, following its seeded 16-frame camera
walk. Scene, LBVH and frame buffers are resident in HBM before timing.

Output: one JSON line with Mpaths/s (pixels x paths/px x frames / s), the
roofline of the dominant kernel (algorithmic bytes per launch / measured
launch time vs 8 TB/s HBM), and a CPU baseline (the oracle, timed on a bounded
sample on the host cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vulkan-restir-pt_amd"))

METRIC = "Mpaths/s @1080p ReSTIR PT; equal-time relMSE, learned vs plain"
PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes per pixel per pass (per-pixel G-buffer / reservoir /
# output traffic, 80-B reservoirs) and per traversal step.
IO_BYTES = {"gbuffer": 28, "path_trace": 104, "temporal": 292, "spatial": 528}
# a 4-wide node visit reads 112 B (4 child boxes + 4 child references); a
# triangle test reads 48 B
NODE_BYTES, TRI_BYTES = 112, 48
PASSES = ["gbuffer", "path_trace", "temporal", "spatial"]
KERNELS = {"gbuffer": "k_gbuffer", "path_trace": "k_gris_path_trace", "temporal": "k_gris_temporal",
           "spatial": "k_gris_spatial"}


# ReSTIR DI passes: per-pixel G-buffer, reservoir (64 B) and output bytes, every neighbour's
# bilinear depth-normal (4 texels), albedo/material and reservoir counted as read.
DI_PASSES = ["gbuffer", "di_path_gen", "di_temporal", "di_spatial"]
DI_IO_BYTES = {"gbuffer": 28, "di_path_gen": 88, "di_temporal": 292, "di_spatial": 1544}
DI_KERNELS = {"gbuffer": "k_gbuffer", "di_path_gen": "k_di_path_gen", "di_temporal": "k_di_temporal",
              "di_spatial": "k_di_spatial"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-frames", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-counters", action="store_true")
    ap.add_argument("--only-kernel-loop", action="store_true", help="time frames only (for profilers)")
    ap.add_argument("--only-overlapped-loop", action="store_true",
                    help="time the headline frames overlapped on two streams only (for profilers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1-GiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--no-learned", action="store_true", help="skip the equal-time relMSE comparison")
    ap.add_argument("--learned-kernel-loop", action="store_true",
                    help="time learned-target frames only (for profilers)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-view leg")
    ap.add_argument("--no-4k", action="store_true", help="skip the 4K full-form leg")
    ap.add_argument("--only-4k", action="store_true", help="run the 4K full-form leg alone (for profilers)")
    ap.add_argument("--guided-kernel-loop", action="store_true",
                    help="time frames with the learned proposal only (for profilers)")
    ap.add_argument("--batched-views", type=int, default=64)
    ap.add_argument("--lvp-frames", type=int, default=64, help="frames of the plain form in the equal-time run")
    ap.add_argument("--ref-frames", type=int, default=4096, help="frames of the converged reference image")
    ap.add_argument("--no-di", action="store_true", help="skip the ReSTIR DI leg")
    ap.add_argument("--no-lbvh", action="store_true", help="skip the LBVH build leg")
    ap.add_argument("--only-lbvh", action="store_true", help="run the LBVH build leg alone (for profilers)")
    ap.add_argument("--only-plain", action="store_true", help="run the 256x256 plain path tracing leg alone")
    ap.add_argument("--learned-study", action="store_true",
                    help="run the paired equal-time study of every learned form alone (long)")
    ap.add_argument("--study-sets", type=int, default=16)
    ap.add_argument("--no-plain", action="store_true", help="skip the 256x256 plain path tracing leg")
    ap.add_argument("--di-kernel-loop", action="store_true", help="time ReSTIR DI frames only (for profilers)")
    ap.add_argument("--no-overlap", action="store_true", help="skip the frames overlapped on two streams")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of exactly 1 GiB, three times: the reference
    point for scaling FETCH_SIZE / WRITE_SIZE counter readings to bytes."""
    import torch

    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}))


def settings(rt, width):
    return rt.settings(shift=2, rr_scale=1.0, temporal=True, spatial=True, cap=20, max_depth=6, neighbors=3,
                       radius=30.0 * width / 1920.0)


class FrameLoop:
    """Runs frames through the C ABI on torch's current stream, with optional
    events between the kernels of each frame."""

    def __init__(self, rt, hs, ds, W, H, s):
        import torch

        self.rt, self.hs, self.W, self.H, self.s = rt, hs, W, H, s
        self.ds = ds
        H_ = rt.host()
        r = C.c_void_p()
        rt.check(H_.rt_renderer_create(ds.handle, W, H, C.byref(r)), "rt_renderer_create")
        self.r = r
        rt.check(H_.rt_renderer_set_settings(r, C.byref(s)), "settings")
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.torch = torch

    def frame(self, k, events=None):
        rt, L, H_ = self.rt, self.rt.core(), self.rt.host()
        cams = self.hs.camera(k % 16, self.W, self.H)
        f = rt.Frame()
        rt.check(H_.rt_renderer_bindings(self.r, C.byref(f)), "bindings")
        ev = events
        if ev: ev[0].record()
        rt.check(L.rt_gbuffer_trace(self.ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId, f.motionVector,
                                    self.stream), "gbuffer")
        if ev: ev[1].record()
        rt.check(L.rt_gris_path_trace(self.ds.handle, cams, C.byref(f), C.byref(self.s), self.stream), "path trace")
        if ev: ev[2].record()
        rt.check(L.rt_gris_temporal(self.ds.handle, cams, C.byref(f), C.byref(self.s), self.stream), "temporal")
        if ev: ev[3].record()
        rt.check(L.rt_gris_spatial(self.ds.handle, cams, C.byref(f), C.byref(self.s), self.stream), "spatial")
        if ev: ev[4].record()
        rt.check(H_.rt_renderer_flip(self.r), "flip")

    def close(self):
        self.rt.host().rt_renderer_destroy(self.r)


def overlapped_frames(rt, hs, ds, W, H, s, args):
    """The headline frames again through rt_pipeline: frame k + 1's G-buffer and path trace on one
    stream beside frame k's temporal and spatial reuse on another (the same frames, computed bit for
    bit as one after the other). HIP gives streams hardware queues as they are created, so a short
    calibration times a few stream pairs and keeps the fastest, then times the path trace's
    persistent launch at a few sizes (fewer workgroups leave CUs to the reuse passes beside it) and
    keeps the fastest. Returns the figures and the last frame's indirect output."""
    import numpy as np
    import torch

    Hh = rt.host()
    p = C.c_void_p()
    rt.check(Hh.rt_pipeline_create(ds.handle, W, H, C.byref(p)), "rt_pipeline_create")
    try:
        rt.check(Hh.rt_pipeline_set_settings(p, C.byref(s)), "settings")
        streams = [torch.cuda.Stream() for _ in range(4)]
        hs_ = [C.c_void_p(x.cuda_stream) for x in streams]

        def run(a, b, first, frames):
            for k in range(first, first + frames):
                rt.check(Hh.rt_pipeline_frame(p, hs.camera(k % 16, W, H), hs_[a], hs_[b]), "pipeline frame")

        def timed(a, b):
            rt.check(Hh.rt_pipeline_reset_reservoirs(p, hs_[a]), "reset")
            run(a, b, 0, 3)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(a, b, 3, 12)
            torch.cuda.synchronize()
            return round((time.perf_counter() - t0) * 1e3 / 12, 3)

        calib = {f"{a}{b}": timed(a, b) for a, b in ((0, 1), (0, 2), (1, 3), (2, 3))}
        pair = min(calib, key=calib.get)
        a, b = int(pair[0]), int(pair[1])
        # one-wave workgroups: 16 per CU is all the path trace's occupancy holds
        resident = torch.cuda.get_device_properties(0).multi_processor_count * 16
        group_calib = {}
        for frac in (16, 12, 10, 8, 6, 4):
            n = resident * frac // 16
            rt.check(Hh.rt_pipeline_set_trace_groups(p, 0 if frac == 16 else n), "trace groups")
            group_calib[n] = timed(a, b)
        groups = min(group_calib, key=group_calib.get)
        rt.check(Hh.rt_pipeline_set_trace_groups(p, 0 if groups == resident else groups), "trace groups")
        rt.check(Hh.rt_pipeline_reset_reservoirs(p, hs_[a]), "reset")
        run(a, b, 0, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(a, b, args.warmup, args.steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out = np.zeros((H, W, 4), np.float32)
        rt.check(Hh.rt_pipeline_read_output(p, 1, out.ctypes.data, hs_[b]), "read output")
    finally:
        torch.cuda.synchronize()
        Hh.rt_pipeline_destroy(p)
    return {"ms_per_step": round(wall * 1e3 / args.steps, 3), "value": round(W * H * args.steps / wall / 1e6, 3),
            "stream_pair_calibration_ms": calib, "streams": pair, "trace_workgroups": groups,
            "trace_workgroups_calibration_ms": group_calib}, out


def count_work(rt, hs, ds, W, H, s, warmup, steps):
    """Traversal work (nodes, triangles, rays) per pass over the timed frames,
    from the counting kernel builds on a fresh replay of the same frames."""
    import numpy as np

    L = rt.core()
    loop = FrameLoop(rt, hs, ds, W, H, s)
    rt.check(L.rt_counters_enable(1), "counters")
    tot = {p: np.zeros(3, np.uint64) for p in PASSES}
    buf = np.zeros(3, np.uint64)
    try:
        for k in range(warmup + steps):
            cams = hs.camera(k % 16, W, H)
            f = rt.Frame()
            rt.check(rt.host().rt_renderer_bindings(loop.r, C.byref(f)), "bindings")
            calls = [
                lambda: L.rt_gbuffer_trace(ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId, f.motionVector,
                                           loop.stream),
                lambda: L.rt_gris_path_trace(ds.handle, cams, C.byref(f), C.byref(s), loop.stream),
                lambda: L.rt_gris_temporal(ds.handle, cams, C.byref(f), C.byref(s), loop.stream),
                lambda: L.rt_gris_spatial(ds.handle, cams, C.byref(f), C.byref(s), loop.stream),
            ]
            for p, call in zip(PASSES, calls):
                rt.check(L.rt_counters_read(buf.ctypes.data, 1), "read")
                rt.check(call(), p)
                rt.check(L.rt_counters_read(buf.ctypes.data, 1), "read")
                if k >= warmup:
                    tot[p] += buf
            rt.check(rt.host().rt_renderer_flip(loop.r), "flip")
    finally:
        L.rt_counters_enable(0)
        loop.close()
    return tot


def cpu_baseline(rt, hs, args):
    """The oracle (scalar C++ restatement) on the host cores over a bounded sample."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc

    W, H = args.width, args.height
    s = settings(rt, W)
    osc = orc.OracleScene(hs.arrays)
    orc.nthreads = lambda: args.cpu_threads
    n = W * H
    this = [orc.reset_reservoirs(n), orc.reset_reservoirs(n)]
    temp = orc.reset_reservoirs(n)
    out = np.zeros((H, W, 4), np.float32)
    dns, ams = [None, None], [None, None]
    cur = 0
    t0 = time.perf_counter()
    for k in range(args.cpu_frames):
        cams = hs.camera(k % 16, W, H)
        dn, am, mv, _ = osc.gbuffer(cams[0])
        dns[cur], ams[cur] = dn, am
        pdn = dns[cur ^ 1] if dns[cur ^ 1] is not None else np.zeros_like(dn)
        pam = ams[cur ^ 1] if ams[cur ^ 1] is not None else np.zeros_like(am)
        osc.gris_path_trace(cams, dn, am, s, this[cur])
        osc.gris_temporal(cams, dn, am, mv, pdn, pam, s, this[cur], this[cur ^ 1], temp)
        osc.gris_spatial(cams, dn, am, s, temp, this[cur], out)
        cur ^= 1
    dt = time.perf_counter() - t0
    return {"value": round(n * args.cpu_frames / dt / 1e6, 4), "unit": "Mpaths/s", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"oracle (scalar C++ restatement of the reference GLSL) on {args.cpu_threads} host threads: "
                      f"{args.cpu_frames} frames of the same pipeline (G-buffer + GRIS path trace + temporal + "
                      f"spatial) at {W}x{H}, same scene, settings and camera walk (frames 0-"
                      f"{args.cpu_frames - 1}), {dt:.1f} s"}


class TorchFrames:
    """Frame images and reservoirs of one view as torch tensors, with the passes
    launched through the C ABI (so that outputs can be accumulated in FP64)."""

    def __init__(self, rt, ds, W, H):
        import torch

        self.rt, self.ds, self.W, self.H, self.torch = rt, ds, W, H, torch
        n = W * H

        def z(nbytes):
            return torch.zeros(nbytes, dtype=torch.uint8, device="cuda")

        self.dn, self.am, self.mv = [z(16 * n), z(16 * n)], [z(8 * n), z(8 * n)], z(4 * n)
        rb = rt.core().rt_gris_reservoir_bytes(W, H)
        self.resv, self.temp = [z(rb), z(rb)], z(rb)
        self.direct = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        self.out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        lb = rt.core().rt_guide_lobe_bytes(W, H)
        self.lobes, self.samples = [z(lb), z(lb)], z(16 * n)
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.cur = 0
        self.reset()

    def reset(self):
        for b in self.resv + [self.temp]:
            self.rt.check(self.rt.core().rt_gris_reservoir_reset(b.data_ptr(), self.W, self.H, self.stream), "reset")
        self.cur = 0

    def frame_struct(self):
        f = self.rt.Frame()
        c, l = self.cur, self.cur ^ 1
        f.width, f.height = self.W, self.H
        f.directOutput, f.indirectOutput = self.direct.data_ptr(), self.out.data_ptr()
        f.depthNormal, f.depthNormalPrev = self.dn[c].data_ptr(), self.dn[l].data_ptr()
        f.albedoMatId, f.albedoMatIdPrev = self.am[c].data_ptr(), self.am[l].data_ptr()
        f.motionVector = self.mv.data_ptr()
        f.grisReservoir, f.grisReservoirPrev = self.resv[c].data_ptr(), self.resv[l].data_ptr()
        f.grisReservoirTemp = self.temp.data_ptr()
        f.guideLobes, f.guideLobesPrev = self.lobes[c].data_ptr(), self.lobes[l].data_ptr()
        f.guideSamples = self.samples.data_ptr()
        return f

    def frame(self, cams, s, mlp=None, events=None, guide=None):
        """G-buffer, [proposal refresh], GRIS path trace, [learned refresh], temporal, spatial, [one training
        step of each network]."""
        rt, L, st = self.rt, self.rt.core(), self.stream
        f = self.frame_struct()
        if events: events[0].record()
        rt.check(L.rt_gbuffer_trace(self.ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId, f.motionVector,
                                    st), "gbuffer")
        if guide is not None:
            rt.check(L.rt_gris_guide_refresh(self.ds.handle, guide.handle, cams, C.byref(f), st), "guide refresh")
        rt.check(L.rt_gris_path_trace(self.ds.handle, cams, C.byref(f), C.byref(s), st), "path trace")
        if mlp is not None:
            rt.check(L.rt_gris_learned_refresh(self.ds.handle, mlp.handle, cams, C.byref(f), st), "refresh")
        rt.check(L.rt_gris_temporal(self.ds.handle, cams, C.byref(f), C.byref(s), st), "temporal")
        rt.check(L.rt_gris_spatial(self.ds.handle, cams, C.byref(f), C.byref(s), st), "spatial")
        if mlp is not None:
            rt.check(L.rt_mlp_train_frame(mlp.handle, None, st), "train")
        if guide is not None:
            rt.check(L.rt_gris_guide_train(self.ds.handle, guide.handle, cams, C.byref(f), None, st), "guide train")
        if events: events[1].record()
        self.cur ^= 1


def static_cams(rt, base, i, seed):
    """Camera pair for frame i of a still view: the first frame's camera with a fresh
    per-frame seed, frameIndex 0 (each frame's output is that frame's estimate; the
    clear flag on frame 0), previous camera = itself."""
    c = (rt.Camera * 2)()
    C.memmove(C.addressof(c), C.addressof(base), C.sizeof(rt.Camera))
    h = (seed * 0x9E3779B1 + i * 0x85EBCA77 + 0x165667B1) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & 0xFFFFFFFF
    c[0].seed = h ^ (h >> 12)
    c[0].frameIndex = 0x80000000 if i == 0 else 0
    for k in range(16):
        c[0].lastProjView[k] = c[0].projView[k]
    C.memmove(C.addressof(c) + C.sizeof(rt.Camera), C.addressof(c), C.sizeof(rt.Camera))
    return c


def mlp_utilisation(rt, ds, fr, s, base, mlp):
    """Matrix-core use of the network kernels on the still 1080p view: one training step (2^16
    samples; algorithmic FLOPs = 3 x 2 x (32x64 + 64x64 + 64x3) per sample for forward, input
    gradient and weight gradient) and one refresh (inference at every surface reconnection vertex),
    each timed over 20 repetitions with HIP events, against the 2.5 PFLOP/s dense FP16 peak."""
    import torch

    L = rt.core()
    for i in range(4):
        fr.frame(static_cams(rt, base, i, 23), s, mlp)
    torch.cuda.synchronize()
    flops_per_sample = 3 * 2 * (32 * 64 + 64 * 64 + 64 * 3)
    cams = static_cams(rt, base, 4, 23)
    f = fr.frame_struct()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    for _ in range(20):
        rt.check(L.rt_mlp_train_frame(mlp.handle, None, fr.stream), "train")
    e[1].record()
    for _ in range(20):
        rt.check(L.rt_gris_learned_refresh(ds.handle, mlp.handle, cams, C.byref(f), fr.stream), "refresh")
    e[2].record()
    torch.cuda.synchronize()
    pos, tgt = mlp.batch()
    train_ms = e[0].elapsed_time(e[1]) / 20
    samples = len(pos)
    tflops = samples * flops_per_sample / (train_ms * 1e-3) / 1e12
    return {"train_step_ms": round(train_ms, 4), "train_samples": samples,
            "train_tflops": round(tflops, 4), "mfma_frac_train": round(tflops / 2500.0, 6),
            "refresh_ms": round(e[1].elapsed_time(e[2]) / 20, 4),
            "note": "gather- and atomic-bound (16 levels x 8 corners per sample); the matrix cores are idle "
                    "most of the time"}


def learned_vs_plain(rt, hs, ds, args):
    """Equal-time relMSE of the learned-target form against the plain (luminance) form
    on a still 1080p view, both judged against a converged reference: the same GRIS
    estimator without reuse (unbiased per frame), args.ref_frames frames, FP64 mean.
    relMSE = mean over pixels of sum_rgb (I - R)^2 / (R^2 + 1e-2)."""
    import torch

    W, H = args.width, args.height
    base = hs.camera(0, W, H)
    fr = TorchFrames(rt, ds, W, H)
    radius = 30.0 * W / 1920.0
    s_plain = settings(rt, W)
    s_ref = rt.settings(shift=2, temporal=False, spatial=False, cap=20, max_depth=6, neighbors=3, radius=radius)
    s_learn = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=6, neighbors=3, radius=radius,
                          learned=True)
    info = ds.info()
    lo, hi = [float(v) for v in info.sceneLo], [float(v) for v in info.sceneHi]
    ext = [max(h - l, 1e-3) for l, h in zip(lo, hi)]
    bounds = ([l - 0.1 * e for l, e in zip(lo, ext)], [h + 0.1 * e for h, e in zip(hi, ext)])

    def new_mlp():
        return rt.Mlp(rt.mlp_config(bounds[0], bounds[1], log2_table=19, seed=3))

    def new_guide():
        return rt.Mlp(rt.mlp_config(bounds[0], bounds[1], log2_table=19, seed=4), guide=True)

    def run(s, frames, seed, mlp=None, guide=None):
        fr.reset()
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(frames)]
        for i in range(frames):
            fr.frame(static_cams(rt, base, i, seed), s, mlp, ev[i], guide)
            acc += fr.out[..., :3].double()
        torch.cuda.synchronize()
        ms = sum(e[0].elapsed_time(e[1]) for e in ev) / frames
        return acc / frames, ms

    t0 = time.perf_counter()
    ref, ref_ms = run(s_ref, args.ref_frames, 7)
    ref_s = time.perf_counter() - t0
    _, plain_ms = run(s_plain, 16, 11)
    _, learn_ms = run(s_learn, 16, 13, new_mlp())
    n_plain = args.lvp_frames
    budget = n_plain * plain_ms
    n_learn = max(1, int(budget // learn_ms))

    def relmse(img):
        return float((((img - ref) ** 2) / (ref ** 2 + 1e-2)).sum(-1).mean())

    def median(v):
        v = sorted(v)
        return (v[(len(v) - 1) // 2] + v[len(v) // 2]) / 2

    # one run's relMSE is dominated by a few fireflies: median (and mean) over 6 independent seed sets;
    # the learned form starts each set from a fresh network, which pays its refresh and training
    # inside the time budget
    sets = (17, 31, 47, 61, 73, 89)
    lv_p, lv_l = [], []
    for seed in sets:
        img_p, plain_ms = run(s_plain, n_plain, seed)
        img_l, learn_ms = run(s_learn, n_learn, seed, new_mlp())
        lv_p.append(relmse(img_p))
        lv_l.append(relmse(img_l))
    rp, rl = median(lv_p), median(lv_l)
    # learned vMF-mixture proposal at equal time: without reuse, with the reference's weights and in the
    # unbiased mode. A single run's relMSE is dominated by a few fireflies (it varies over 25x between
    # seed sets), so each figure is the median (and the mean) over 6 independent seed sets; the network
    # is warm (trained on 128 earlier frames of the view with other seeds) and keeps training during the
    # measured frames.
    guided = {}
    for mode, unb, reuse in (("no_reuse", False, False), ("reference_weights", False, True), ("unbiased", True, True)):
        kw = dict(shift=2, temporal=reuse, spatial=reuse, cap=20, max_depth=6, neighbors=3, radius=radius,
                  unbiased=unb)
        s_pl, s_gu = rt.settings(**kw), rt.settings(guided=True, **kw)
        warm = new_guide()
        _, pl_ms = run(s_pl, 16, 11)
        run(s_gu, 128, 29, None, warm)
        _, gu_ms = run(s_gu, 16, 13, None, warm)
        n_pl = args.lvp_frames
        n_gu = max(1, int(n_pl * pl_ms // gu_ms))
        r_pl, r_gu = [], []
        for seed in sets:
            img_pl, _ = run(s_pl, n_pl, seed)
            img_gu, _ = run(s_gu, n_gu, seed, None, warm)
            r_pl.append(relmse(img_pl))
            r_gu.append(relmse(img_gu))
        m_pl, m_gu = sum(r_pl) / len(sets), sum(r_gu) / len(sets)
        med_pl, med_gu = median(r_pl), median(r_gu)
        guided[mode] = {"relmse_plain_median": round(med_pl, 6), "relmse_guided_median": round(med_gu, 6),
                        "plain_over_guided_median": round(med_pl / med_gu, 4) if med_gu > 0 else None,
                        "relmse_plain_mean": round(m_pl, 6), "relmse_guided_mean": round(m_gu, 6),
                        "per_set": {"plain": [round(x, 4) for x in r_pl], "guided": [round(x, 4) for x in r_gu]},
                        "frames": {"plain": n_pl, "guided": n_gu},
                        "ms_per_frame": {"plain": round(pl_ms, 3), "guided": round(gu_ms, 3)}}
    mlp_perf = mlp_utilisation(rt, ds, fr, s_learn, base, new_mlp())
    return {"relmse_plain": round(rp, 6), "relmse_learned": round(rl, 6),
            "plain_over_learned": round(rp / rl, 4) if rl > 0 else None,
            "statistic": "median over 6 seed sets",
            "per_set": {"plain": [round(x, 4) for x in lv_p], "learned": [round(x, 4) for x in lv_l]},
            "guided_proposal": guided,
            "frames": {"plain": n_plain, "learned": n_learn}, "ms_per_frame": {"plain": round(plain_ms, 3),
                                                                                "learned": round(learn_ms, 3)},
            "equal_time_ms": round(budget, 1), "mlp": mlp_perf,
            "reference": f"GRIS without reuse, {args.ref_frames} frames ({ref_ms:.2f} ms each, {ref_s:.1f} s), FP64 mean",
            "view": f"terrain camera frame 0 held still, {W}x{H}, per-frame seeds; learned: hash grid 16x2x2^19 FP16 "
                    f"+ MLP 2x64 FP16 (MFMA), trained online each frame on 2^16 reconnection samples, Adam lr 1e-2"}


def learned_study(rt, hs, ds, args):
    """Equal-time relMSE of every learned form against its plain counterpart, paired over
    args.study_sets independent seed sets on the still view of learned_vs_plain (reference: the
    GRIS estimator without reuse, args.ref_frames frames, FP64 mean). For each mode (no reuse, the
    reference's resampling weights, unbiased reuse) the plain form renders args.lvp_frames frames
    and each learned form (learned target, learned proposal, both) as many frames as fit the same
    GPU time, its networks warm (trained on 128 earlier frames of the view with other seeds) and
    training on. Per form: the geometric mean over sets of relMSE(plain) / relMSE(learned) with a
    95 % t-interval on the mean log ratio, the median ratio, and the sets where the learned form is
    better. A set's relMSE is dominated by a few fireflies, which is why the comparison is paired
    and on the log scale."""
    import math

    import torch

    W, H = args.width, args.height
    base = hs.camera(0, W, H)
    fr = TorchFrames(rt, ds, W, H)
    radius = 30.0 * W / 1920.0
    info = ds.info()
    lo, hi = [float(v) for v in info.sceneLo], [float(v) for v in info.sceneHi]
    ext = [max(h - l, 1e-3) for l, h in zip(lo, hi)]
    bounds = ([l - 0.1 * e for l, e in zip(lo, ext)], [h + 0.1 * e for h, e in zip(hi, ext)])

    def run(s, frames, seed, mlp=None, guide=None):
        fr.reset()
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(frames)]
        for i in range(frames):
            fr.frame(static_cams(rt, base, i, seed), s, mlp, ev[i], guide)
            acc += fr.out[..., :3].double()
        torch.cuda.synchronize()
        return acc / frames, sum(e[0].elapsed_time(e[1]) for e in ev) / frames

    s_ref = rt.settings(shift=2, temporal=False, spatial=False, cap=20, max_depth=6, neighbors=3, radius=radius)
    ref, _ = run(s_ref, args.ref_frames, 7)

    def relmse(img):
        return float((((img - ref) ** 2) / (ref ** 2 + 1e-2)).sum(-1).mean())

    # two-sided 95 % Student t quantiles by degrees of freedom
    tq = {5: 2.571, 7: 2.365, 9: 2.262, 11: 2.201, 15: 2.131, 19: 2.093, 23: 2.069, 31: 2.040}
    sets = [1000 + 37 * k for k in range(args.study_sets)]
    t95 = tq.get(len(sets) - 1, 2.0)
    out = {"sets": len(sets), "frames_plain": args.lvp_frames}
    for mode, unb, reuse in (("no_reuse", False, False), ("reference_weights", False, True), ("unbiased", True, True)):
        kw = dict(shift=2, temporal=reuse, spatial=reuse, cap=20, max_depth=6, neighbors=3, radius=radius, unbiased=unb)
        forms = {"plain": rt.settings(**kw), "target": rt.settings(learned=True, **kw),
                 "proposal": rt.settings(guided=True, **kw), "both": rt.settings(learned=True, guided=True, **kw)}
        if not reuse:
            del forms["target"], forms["both"]   # the learned target only enters the reuse passes
        nets = {}
        ms = {}
        for name, s in forms.items():
            mlp = rt.Mlp(rt.mlp_config(bounds[0], bounds[1], log2_table=19, seed=3)) if s.targetFunction else None
            guide = (rt.Mlp(rt.mlp_config(bounds[0], bounds[1], log2_table=19, seed=4), guide=True)
                     if s.guidedProposal else None)
            if mlp or guide:
                run(s, 128, 29, mlp, guide)
            _, ms[name] = run(s, 16, 13, mlp, guide)
            nets[name] = (mlp, guide)
        budget = args.lvp_frames * ms["plain"]
        frames = {k: max(1, int(budget // v)) if k != "plain" else args.lvp_frames for k, v in ms.items()}
        err = {k: [] for k in forms}
        for seed in sets:
            for name, s in forms.items():
                img, _ = run(s, frames[name], seed, *nets[name])
                err[name].append(relmse(img))
        res = {"ms_per_frame": {k: round(v, 3) for k, v in ms.items()}, "frames": frames,
               "relmse_median": {k: round(sorted(v)[len(v) // 2], 5) for k, v in err.items()},
               "per_set": {k: [round(x, 5) for x in v] for k, v in err.items()}}
        for name in forms:
            if name == "plain":
                continue
            lr = [math.log(p / q) for p, q in zip(err["plain"], err[name]) if p > 0 and q > 0]
            m = sum(lr) / len(lr)
            sd = math.sqrt(sum((x - m) ** 2 for x in lr) / (len(lr) - 1)) if len(lr) > 1 else float("nan")
            half = t95 * sd / math.sqrt(len(lr))
            ratios = sorted(math.exp(x) for x in lr)
            res[f"plain_over_{name}"] = {"geomean": round(math.exp(m), 4),
                                         "ci95": [round(math.exp(m - half), 4), round(math.exp(m + half), 4)],
                                         "median": round(ratios[len(ratios) // 2], 4),
                                         "better_in": f"{sum(1 for x in lr if x > 0)} of {len(lr)}"}
        out[mode] = res
        progress(f"learned study: {mode} done")
    return out


def learned_kernel_loop(rt, hs, ds, args):
    """Learned-target frames of the benchmark camera walk (path trace, refresh, temporal,
    spatial, one training step) with per-pass HIP events, for profiling."""
    import torch

    W, H = args.width, args.height
    fr = TorchFrames(rt, ds, W, H)
    s = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=6, neighbors=3,
                    radius=30.0 * W / 1920.0, learned=True)
    info = ds.info()
    lo, hi = [float(v) for v in info.sceneLo], [float(v) for v in info.sceneHi]
    ext = [max(h - l, 1e-3) for l, h in zip(lo, hi)]
    mlp = rt.Mlp(rt.mlp_config([l - 0.1 * e for l, e in zip(lo, ext)], [h + 0.1 * e for h, e in zip(hi, ext)],
                               log2_table=19, seed=3))
    for k in range(args.warmup):
        fr.frame(hs.camera(k % 16, W, H), s, mlp)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        fr.frame(hs.camera((args.warmup + k) % 16, W, H), s, mlp, ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"learned_frames": args.steps, "ms_per_frame": round(wall * 1e3 / args.steps, 3),
            "gpu_ms_per_frame": round(sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps, 3),
            "mpaths_per_s": round(W * H * args.steps / wall / 1e6, 3)}


def guided_kernel_loop(rt, hs, ds, args):
    """Frames with the learned proposal on the benchmark camera walk (G-buffer, proposal refresh, path
    trace, temporal, spatial, one proposal training step), unbiased reuse, with HIP events around the
    proposal's own work, for profiling."""
    import torch

    W, H = args.width, args.height
    fr = TorchFrames(rt, ds, W, H)
    s = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=6, neighbors=3,
                    radius=30.0 * W / 1920.0, unbiased=True, guided=True)
    info = ds.info()
    lo, hi = [float(v) for v in info.sceneLo], [float(v) for v in info.sceneHi]
    ext = [max(h - l, 1e-3) for l, h in zip(lo, hi)]
    guide = rt.Mlp(rt.mlp_config([l - 0.1 * e for l, e in zip(lo, ext)], [h + 0.1 * e for h, e in zip(hi, ext)],
                                 log2_table=19, seed=4), guide=True)
    for k in range(args.warmup):
        fr.frame(hs.camera(k % 16, W, H), s, None, None, guide)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        fr.frame(hs.camera((args.warmup + k) % 16, W, H), s, None, ev[k], guide)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"guided_frames": args.steps, "ms_per_frame": round(wall * 1e3 / args.steps, 3),
            "gpu_ms_per_frame": round(sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps, 3)}


def batched_views(rt, args, hs1, ds1):
    """Batched plain ReSTIR PT: args.batched_views views over four scenes (~20k, ~225k, ~1.1M and ~3.0M
    triangles), square resolutions log-uniform in [128, 1024] (multiples of 8), max depth uniform in
    {3..8}, each view on its own frame of its scene's camera walk; one rt_batch_frame per frame after
    the autotuner has picked the persistent grid and tile order. Also times the same frames launched
    view by view."""
    import numpy as np
    import torch

    g = np.random.default_rng(2)
    hosts = [rt.HostScene(0, 2), rt.HostScene(3, 2), hs1, rt.HostScene(4, 2)]
    devs = [rt.DeviceScene(hosts[0].arrays), rt.DeviceScene(hosts[1].arrays), ds1, rt.DeviceScene(hosts[3].arrays)]
    tris = [int(d.info().numPrimitives) for d in devs]
    specs = []
    for i in range(args.batched_views):
        size = int(np.exp(g.uniform(np.log(128), np.log(1024)))) // 8 * 8
        specs.append((i % 4, size, int(g.integers(3, 9)), int(g.integers(0, 64))))
    frames = [TorchFrames(rt, devs[si], n, n) for si, n, _, _ in specs]
    sets = [rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=d, neighbors=3, radius=20.0)
            for _, _, d, _ in specs]
    pixels = sum(n * n for _, n, _, _ in specs)
    L = rt.core()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def views_for(k):
        vs = (rt.View * len(specs))()
        for i, (si, n, _, f0) in enumerate(specs):
            cams = hosts[si].camera((f0 + k) % 64, n, n)
            v = vs[i]
            v.scene = devs[si].handle.value
            C.memmove(C.addressof(v.cams), C.addressof(cams), C.sizeof(rt.Camera) * 2)
            v.frame = frames[i].frame_struct()
            v.settings = sets[i]
        return vs

    def flip():
        for fr in frames:
            fr.cur ^= 1

    b = C.c_void_p()
    rt.check(L.rt_batch_create(len(specs), C.byref(b)), "batch")
    try:
        best, tune_ms = rt.BatchTuning(), C.c_float()
        # tuned on frame 1 of each view's walk: frame 0's camera carries the first-frame flag, under which
        # temporal reuse does nothing, so frames tuned on it would not represent the timed ones
        rt.check(L.rt_batch_autotune(b, views_for(1), len(specs), 2, stream, C.byref(best), C.byref(tune_ms)), "tune")
        for fr in frames:
            fr.reset()
        for k in range(2):
            rt.check(L.rt_batch_frame(b, views_for(k), len(specs), stream), "frame")
            flip()
        torch.cuda.synchronize()
        steps = 8
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        host_s = 0.0
        t0 = time.perf_counter()
        e0.record()
        for k in range(steps):
            h0 = time.perf_counter()
            rt.check(L.rt_batch_frame(b, views_for(2 + k), len(specs), stream), "frame")
            host_s += time.perf_counter() - h0
            flip()
        e1.record()
        torch.cuda.synchronize()
        batched_ms = (time.perf_counter() - t0) * 1e3 / steps
        gpu_ms = e0.elapsed_time(e1) / steps

        def timed_gpu(tuning, first, n):
            rt.check(L.rt_batch_set_tuning(b, C.byref(tuning)), "tuning")
            rt.check(L.rt_batch_frame(b, views_for(first), len(specs), stream), "frame")  # warm-up
            flip()
            a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for k in range(n):
                rt.check(L.rt_batch_frame(b, views_for(first + 1 + k), len(specs), stream), "frame")
                flip()
            z.record()
            torch.cuda.synchronize()
            return a.elapsed_time(z) / n

        # both path-trace variants at the tuned settings: the persistent kernel and the
        # wavefront variant with material-sorted shading (the tuner keeps the faster)
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        variant_ms = {}
        for shading in (0, 1):
            t = rt.BatchTuning.from_buffer_copy(bytes(best))
            if t.shading != shading:
                t.shading, t.persistentGroups = shading, (4 * cus if shading else 0)
            variant_ms[("persistent", "wavefront_material_sorted")[shading]] = round(timed_gpu(t, 20, 4), 3)
        rt.check(L.rt_batch_set_tuning(b, C.byref(best)), "tuning")
        # the same frames launched view by view (4 launches per view)
        t0 = time.perf_counter()
        for k in range(2):
            vs = views_for(40 + k)
            for i in range(len(specs)):
                v = vs[i]
                cams = v.cams
                rt.check(L.rt_gbuffer_trace(v.scene, C.byref(cams[0]), v.frame.depthNormal, v.frame.albedoMatId,
                                            v.frame.motionVector, stream), "gbuffer")
                for fn in (L.rt_gris_path_trace, L.rt_gris_temporal, L.rt_gris_spatial):
                    rt.check(fn(v.scene, cams, C.byref(v.frame), C.byref(v.settings), stream), "pass")
            flip()
        torch.cuda.synchronize()
        single_ms = (time.perf_counter() - t0) * 1e3 / 2
    finally:
        L.rt_batch_destroy(b)
    return {"views": len(specs), "pixels_per_frame": pixels, "scene_triangles": tris,
            "ms_per_frame": round(batched_ms, 3), "mpaths_per_s": round(pixels / batched_ms / 1e3, 2),
            "gpu_ms_per_frame": round(gpu_ms, 3), "host_ms_per_frame": round(host_s * 1e3 / steps, 3),
            "path_trace_variants_gpu_ms_per_frame": variant_ms,
            "view_by_view_ms_per_frame": round(single_ms, 3),
            "tuning": {"persistent_groups": int(best.persistentGroups), "xcd_chunk_tiles": int(best.xcdSwizzle),
                       "waves": int(best.waves), "queue_tile": ["8x8", "16x4", "64x1"][best.queueTile],
                       "graph": int(best.graph), "shading": ["persistent", "wavefront_material_sorted"][best.shading],
                       "tuned_ms": round(float(tune_ms.value), 3)},
            "workload": "batched plain ReSTIR PT (hybrid shift, temporal + spatial, 3 neighbours in 20 px, "
                        "M-cap 20), square views log-uniform in [128, 1024] px, depth U{3..8}, scenes "
                        "round robin over 4 procedural scenes"}


def full_form_4k(rt, args):
    """The full-form workload at 3840x2160 on the ~4.8M-triangle scene (1024x1024-cell fBm heightfield
    + 2000 instanced icospheres, 64 emissive triangles, GGX roughness U[0.02, 1]), depth 8, hybrid
    shift, temporal + spatial: per frame one batched frame (rt_batch_frame) of the 4K view with the
    learned target (refresh after path tracing, one training step after spatial reuse), the learned
    vMF-mixture proposal (refresh after the G-buffer, one training step after spatial reuse) and
    unbiased reuse, alongside 16 extra 512x512 views of the same scene (luminance target, BSDF
    sampling): one persistent path-trace launch over all 17 views, G-buffer and reuse launches per
    form."""
    import numpy as np
    import torch

    hs = rt.HostScene(5, 4)
    ds = rt.DeviceScene(hs.arrays)
    W, H = 3840, 2160
    fr = TorchFrames(rt, ds, W, H)
    s4k = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=8, neighbors=3, radius=30.0,
                      learned=True, unbiased=True, guided=True)
    info = ds.info()
    lo, hi = [float(v) for v in info.sceneLo], [float(v) for v in info.sceneHi]
    ext = [max(h - l, 1e-3) for l, h in zip(lo, hi)]
    box = ([l - 0.1 * e for l, e in zip(lo, ext)], [h + 0.1 * e for h, e in zip(hi, ext)])
    mlp = rt.Mlp(rt.mlp_config(box[0], box[1], log2_table=19, seed=4))
    guide = rt.Mlp(rt.mlp_config(box[0], box[1], log2_table=19, seed=5), guide=True)
    n_extra, size = 16, 512
    extra = [TorchFrames(rt, ds, size, size) for _ in range(n_extra)]
    s_extra = rt.settings(shift=2, temporal=True, spatial=True, cap=20, max_depth=8, neighbors=3, radius=30.0)
    offsets = np.random.default_rng(4).integers(0, 64, n_extra)
    L = rt.core()
    stream = fr.stream
    n_views = 1 + n_extra

    def views_for(k):
        vs = (rt.View * n_views)()
        for i in range(n_views):
            v = vs[i]
            v.scene = ds.handle.value
            if i == 0:
                cams = hs.camera(k % 64, W, H)
                v.frame, v.settings = fr.frame_struct(), s4k
                v.target, v.guide = mlp.handle.value, guide.handle.value
            else:
                cams = hs.camera(int(offsets[i - 1] + k) % 64, size, size)
                v.frame, v.settings = extra[i - 1].frame_struct(), s_extra
            C.memmove(C.addressof(v.cams), C.addressof(cams), C.sizeof(rt.Camera) * 2)
        return vs

    b = C.c_void_p()
    rt.check(L.rt_batch_create(n_views, C.byref(b)), "batch")
    try:
        # tuned on these views (the tuning frames also train the 4K view's networks: warm-up)
        best, tune_ms = rt.BatchTuning(), C.c_float()
        # frame 1 of the walk (frame 0's first-frame flag would switch temporal reuse off while tuning)
        rt.check(L.rt_batch_autotune(b, views_for(1), n_views, 2, stream, C.byref(best), C.byref(tune_ms)), "tune")
        for e in [fr] + extra:
            e.reset()

        def frame(k, ev=None):
            if ev is not None:
                ev[0].record()
            rt.check(L.rt_batch_frame(b, views_for(k), n_views, stream), "batch frame")
            if ev is not None:
                ev[1].record()
            for e in [fr] + extra:
                e.cur ^= 1

        for k in range(2):
            frame(k)
        torch.cuda.synchronize()
        steps = 8
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            frame(2 + k, ev[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        L.rt_batch_destroy(b)
    pixels = W * H + n_extra * size * size
    batch_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    return {"ms_per_frame": round(wall * 1e3 / steps, 3), "mpaths_per_s": round(pixels * steps / wall / 1e6, 2),
            "batched_frame_gpu_ms": round(batch_ms, 3),
            "pixels_per_frame": pixels, "triangles": int(info.numPrimitives), "emissive": int(hs.arrays.numLights),
            "tuning": {"persistent_groups": int(best.persistentGroups), "xcd_chunk_tiles": int(best.xcdSwizzle),
                       "waves": int(best.waves), "queue_tile": ["8x8", "16x4", "64x1"][best.queueTile],
                       "graph": int(best.graph), "shading": ["persistent", "wavefront_material_sorted"][best.shading],
                       "tuned_ms": round(float(tune_ms.value), 3)},
            "workload": "one rt_batch_frame of 17 views: a 3840x2160 view with the learned target (hash grid "
                        "16x2x2^19 + MLP 2x64) and the learned 8-lobe vMF-mixture proposal (second network, 32 "
                        "outputs), both trained each frame, unbiased reuse; + 16 x 512x512 views (luminance "
                        "target, BSDF sampling); ~4.8M-triangle scene, 64 emissive, depth 8, hybrid shift, "
                        "temporal + spatial (3 neighbours, 30 px, M-cap 20)"}


def lbvh_bytes(n):
    """Algorithmic HBM bytes of one binary LBVH build over n triangles: the triangles read once
    (48 B) and the (Morton code, primitive) pairs written once (12 B); each of the 8 LSD radix
    passes reads and writes the pairs (24 B); topology and refit write and read the 2n - 1 node
    slots (64 B per node: the (2N-1) x 64 B of the build's roofline)."""
    return 48 * n + 12 * n + 8 * 24 * n + 64 * (2 * n - 1)


def lbvh_build(rt, scenes, reps=3):
    """GPU time of the LBVH build (rt_device_scene_create: world triangles, bounds, Morton codes, radix
    sort, Karras topology, refit, 4-wide collapse) per scene, best of reps builds, and the roofline
    of its binary part against 8 TB/s."""
    out = {}
    for label, hs in scenes:
        best_bin, best_tot, n = None, None, 0
        for _ in range(reps):
            ds = rt.DeviceScene(hs.arrays)
            info = ds.info()
            n = int(info.numPrimitives)
            b, t = float(info.bvhBinaryMs), float(info.bvhBuildMs)
            best_bin = b if best_bin is None else min(best_bin, b)
            best_tot = t if best_tot is None else min(best_tot, t)
            depth, depth4 = int(info.bvhDepth), int(info.bvh4Depth)
            del ds
        byts = lbvh_bytes(n)
        gbs = byts / (best_bin * 1e-3) / 1e9
        out[label] = {"primitives": n, "binary_ms": round(best_bin, 3), "total_ms": round(best_tot, 3),
                      "bvh_depth": depth, "bvh4_depth": depth4,
                      "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                   "frac": round(gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": byts}}
    return out


def pt_plain_c0(rt, args):
    """Plain path tracing with NEE+MIS (di_naive.comp + gi_naive.comp through rt_pt_plain) at 256x256,
    1 spp, depth 4, on the box room (12 quads + 64 icospheres of 320 triangles, one ceiling quad
    light): per frame a G-buffer and the plain path tracer (direct + indirect), following the
    scene's camera path; Mpaths/s, Mrays/s from the counting builds, and the oracle on the host
    cores over the same frames."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc

    hs = rt.HostScene(0, 0)
    ds = rt.DeviceScene(hs.arrays)
    W = H = 256
    depth = 4
    fr = TorchFrames(rt, ds, W, H)
    L = rt.core()

    def frame(k):
        cams = hs.camera(k % 16, W, H)
        f = fr.frame_struct()
        rt.check(L.rt_gbuffer_trace(ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId, f.motionVector,
                                    fr.stream), "gbuffer")
        rt.check(L.rt_pt_plain(ds.handle, cams, C.byref(f), depth, 3, fr.stream), "pt plain")
        fr.cur ^= 1

    for k in range(5):
        frame(k)
    torch.cuda.synchronize()
    steps = 400
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for k in range(steps):
        frame(5 + k)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1) / steps
    buf = np.zeros(3, np.uint64)
    rt.check(L.rt_counters_enable(1), "counters")
    try:
        rt.check(L.rt_counters_read(buf.ctypes.data, 1), "read")
        for k in range(16):
            frame(5 + k)
        torch.cuda.synchronize()
        rt.check(L.rt_counters_read(buf.ctypes.data, 1), "read")
    finally:
        L.rt_counters_enable(0)
    rays_per_frame = int(buf[2]) / 16
    # the oracle on the host cores, same frames, bounded to about 10 s
    osc = orc.OracleScene(hs.arrays)
    orc.nthreads = lambda: args.cpu_threads
    t1 = time.perf_counter()
    frames = 0
    while frames < 200 and (frames < 4 or time.perf_counter() - t1 < 10.0):
        cams = hs.camera((5 + frames) % 16, W, H)
        dn, am, _, _ = osc.gbuffer(cams[0])
        osc.pt_plain(cams[0], dn, am, depth)
        frames += 1
    cpu_s = time.perf_counter() - t1
    return {"workload": "plain PT with NEE+MIS (di_naive + gi_naive), 256x256, 1 spp, depth 4, box room with 64 "
                        "icospheres (%d triangles), G-buffer + plain PT per frame" % int(ds.info().numPrimitives),
            "ms_per_frame": round(wall * 1e3 / steps, 4), "gpu_ms_per_frame": round(gpu_ms, 4),
            "mpaths_per_s": round(W * H * steps / wall / 1e6, 2),
            "mrays_per_s": round(rays_per_frame * steps / wall / 1e6, 1), "rays_per_frame": int(rays_per_frame),
            "cpu_baseline": {"value": round(W * H * frames / cpu_s / 1e6, 4), "unit": "Mpaths/s",
                             "cores": args.cpu_threads, "kind": "port",
                             "sample": f"oracle G-buffer + pt_plain on {args.cpu_threads} host threads, {frames} "
                                       f"frames of the same camera path at 256x256, {cpu_s:.1f} s"}}


def restir_di(rt, hs, ds, args, kernel_loop=False):
    """ReSTIR DI (the reference's TestReSTIR: path generation, temporal reuse, spatial reuse with
    10 neighbours in 20 px and M-cap 40) at the bench resolution on the same terrain and camera walk,
    with the reference's default settings {Reconnection, Light, temporal off, spatial on} and with
    temporal reuse on and both sample types. Per-pass times from HIP events on the launch stream;
    traversal work from the counting kernel builds on a replay of the same frames."""
    import numpy as np
    import torch

    W, H = args.width, args.height
    L, H_ = rt.core(), rt.host()
    r = C.c_void_p()
    rt.check(H_.rt_renderer_create(ds.handle, W, H, C.byref(r)), "rt_renderer_create")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    variants = {"reference_defaults": rt.di_settings(shift=0, sample=0, temporal=False, spatial=True),
                "temporal_spatial_both": rt.di_settings(shift=0, sample=2, temporal=True, spatial=True)}
    if kernel_loop:
        variants = {"reference_defaults": variants["reference_defaults"]}
    res = {}
    cnt = np.zeros(3, np.uint64)

    def frame(k, s, ev=None, work=None):
        cams = hs.camera(k % 16, W, H)
        f, d = rt.Frame(), rt.DIReservoirs()
        rt.check(H_.rt_renderer_bindings(r, C.byref(f)), "bindings")
        rt.check(H_.rt_renderer_di_bindings(r, C.byref(d)), "di bindings")
        calls = [lambda: L.rt_gbuffer_trace(ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId,
                                            f.motionVector, stream),
                 lambda: L.rt_di_path_gen(ds.handle, cams, C.byref(f), C.byref(d), C.byref(s), stream),
                 lambda: L.rt_di_temporal(ds.handle, cams, C.byref(f), C.byref(d), C.byref(s), stream),
                 lambda: L.rt_di_spatial(ds.handle, cams, C.byref(f), C.byref(d), C.byref(s), stream)]
        for i, (p, call) in enumerate(zip(DI_PASSES, calls)):
            if ev: ev[i].record()
            if work is not None: rt.check(L.rt_counters_read(cnt.ctypes.data, 1), "read")
            rt.check(call(), p)
            if work is not None:
                rt.check(L.rt_counters_read(cnt.ctypes.data, 1), "read")
                work[p] += cnt
        if ev: ev[4].record()
        rt.check(H_.rt_renderer_flip(r), "flip")

    try:
        for name, s in variants.items():
            rt.check(H_.rt_renderer_reset_reservoirs(r, stream), "reset")
            for k in range(args.warmup):
                frame(k, s)
            torch.cuda.synchronize()
            ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
            t0 = time.perf_counter()
            for k in range(args.steps):
                frame(args.warmup + k, s, None if kernel_loop else ev[k])
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            v = {"ms_per_frame": round(wall * 1e3 / args.steps, 3),
                 "mpixels_per_s": round(W * H * args.steps / wall / 1e6, 2)}
            if not kernel_loop:
                pass_ms = {p: float(np.mean([ev[k][i].elapsed_time(ev[k][i + 1]) for k in range(args.steps)]))
                           for i, p in enumerate(DI_PASSES)}
                v["passes_ms"] = {p: round(x, 3) for p, x in pass_ms.items()}
                work = {p: np.zeros(3, np.uint64) for p in DI_PASSES}
                rt.check(L.rt_counters_enable(1), "counters")
                try:
                    for k in range(args.steps):
                        frame(args.warmup + k, s, work=work)
                finally:
                    L.rt_counters_enable(0)
                rays = sum(int(work[p][2]) for p in DI_PASSES)
                v["mrays_per_s"] = round(rays / (sum(pass_ms.values()) * 1e-3 * args.steps) / 1e6, 1)
                dom = max(DI_PASSES[1:], key=lambda p: pass_ms[p])
                per_launch = (DI_IO_BYTES[dom] * W * H + (NODE_BYTES * int(work[dom][0]) +
                                                         TRI_BYTES * int(work[dom][1])) / args.steps)
                achieved = per_launch / (pass_ms[dom] * 1e-3) / 1e9
                v["roofline"] = {"bound": "hbm", "kernel": DI_KERNELS[dom], "achieved": round(achieved, 1),
                                 "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                                 "algorithmic_bytes_per_launch": int(per_launch), "launch_ms": round(pass_ms[dom], 3)}
                v["work_per_frame"] = {p: {"rays": int(work[p][2]) // args.steps,
                                           "nodes": int(work[p][0]) // args.steps,
                                           "tris": int(work[p][1]) // args.steps} for p in DI_PASSES}
            res[name] = v
    finally:
        H_.rt_renderer_destroy(r)
    res["workload"] = (f"ReSTIR DI (TestReSTIR) frames: G-buffer + path generation + temporal + spatial "
                       f"(10 neighbours, 20 px, M-cap 40, reconnection shift), {W}x{H}, same terrain and camera walk")
    return res


def full_image(rt, hs, ds, args):
    """The whole image as the reference composes it with both resampled methods: ReSTIR DI for the
    direct light (the reference's defaults) and ReSTIR PT for the indirect light, on the headline's
    view, settings and camera walk. Frames one after the other (rt_renderer_frame_methods(2, 3)) and
    overlapped on two streams (rt_pipeline with the direct method 2); the last frame's direct and
    indirect outputs must agree bit for bit."""
    import numpy as np
    import torch

    W, H = args.width, args.height
    Hh = rt.host()
    s = settings(rt, W)
    di = rt.di_settings()
    frames = args.warmup + args.steps
    outs = {}

    def read(fn, handle, which, stream):
        o = np.zeros((H, W, 4), np.float32)
        rt.check(fn(handle, which, o.ctypes.data, stream), "read output")
        return o

    r = C.c_void_p()
    rt.check(Hh.rt_renderer_create(ds.handle, W, H, C.byref(r)), "rt_renderer_create")
    try:
        rt.check(Hh.rt_renderer_set_settings(r, C.byref(s)), "settings")
        rt.check(Hh.rt_renderer_set_di_settings(r, C.byref(di)), "di settings")
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for k in range(args.warmup):
            rt.check(Hh.rt_renderer_frame_methods(r, hs.camera(k % 16, W, H), 2, 3, st), "frame")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.warmup, frames):
            rt.check(Hh.rt_renderer_frame_methods(r, hs.camera(k % 16, W, H), 2, 3, st), "frame")
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        outs["seq"] = (read(Hh.rt_renderer_read_output, r, 0, st), read(Hh.rt_renderer_read_output, r, 1, st))
    finally:
        Hh.rt_renderer_destroy(r)
    p = C.c_void_p()
    rt.check(Hh.rt_pipeline_create(ds.handle, W, H, C.byref(p)), "rt_pipeline_create")
    try:
        rt.check(Hh.rt_pipeline_set_settings(p, C.byref(s)), "settings")
        rt.check(Hh.rt_pipeline_set_di_settings(p, C.byref(di)), "di settings")
        rt.check(Hh.rt_pipeline_set_direct_method(p, 2), "direct method")
        streams = [torch.cuda.Stream() for _ in range(3)]
        ta, tb, tc = (C.c_void_p(x.cuda_stream) for x in streams)
        # ReSTIR DI's reuse on a third stream, beside the path trace and the GRIS reuse
        rt.check(Hh.rt_pipeline_set_direct_stream(p, tc), "direct stream")
        # the path trace's launch size, timed over a few frames at each cap as for the headline
        resident = torch.cuda.get_device_properties(0).multi_processor_count * 16
        cap_ms = {}
        for frac in (16, 12, 10, 8):
            n = resident * frac // 16
            rt.check(Hh.rt_pipeline_set_trace_groups(p, 0 if frac == 16 else n), "trace groups")
            rt.check(Hh.rt_pipeline_reset_reservoirs(p, ta), "reset")
            for k in range(3):
                rt.check(Hh.rt_pipeline_frame(p, hs.camera(k % 16, W, H), ta, tb), "frame")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(3, 11):
                rt.check(Hh.rt_pipeline_frame(p, hs.camera(k % 16, W, H), ta, tb), "frame")
            torch.cuda.synchronize()
            cap_ms[n] = round((time.perf_counter() - t0) * 1e3 / 8, 3)
        groups = min(cap_ms, key=cap_ms.get)
        rt.check(Hh.rt_pipeline_set_trace_groups(p, 0 if groups == resident else groups), "trace groups")
        rt.check(Hh.rt_pipeline_reset_reservoirs(p, ta), "reset")
        for k in range(args.warmup):
            rt.check(Hh.rt_pipeline_frame(p, hs.camera(k % 16, W, H), ta, tb), "frame")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.warmup, frames):
            rt.check(Hh.rt_pipeline_frame(p, hs.camera(k % 16, W, H), ta, tb), "frame")
        torch.cuda.synchronize()
        ov_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        outs["ov"] = (read(Hh.rt_pipeline_read_output, p, 0, tc), read(Hh.rt_pipeline_read_output, p, 1, tb))
    finally:
        torch.cuda.synchronize()
        Hh.rt_pipeline_destroy(p)
    same = all(((a == b) | (np.isnan(a) & np.isnan(b))).all() for a, b in zip(outs["seq"], outs["ov"]))
    return {"ms_per_frame_sequential": round(seq_ms, 3), "ms_per_frame_overlapped": round(ov_ms, 3),
            "mpixels_per_s_overlapped": round(W * H / ov_ms / 1e3, 2), "identical_to_sequential": bool(same),
            "trace_workgroups": groups, "trace_workgroups_calibration_ms": cap_ms,
            "workload": f"ReSTIR DI direct (reference defaults) + ReSTIR PT indirect (headline settings), "
                        f"{W}x{H}, terrain camera walk; overlapped: rt_pipeline on three streams (trace, "
                        f"GRIS reuse, DI reuse)"}


def restir_gi(rt, hs, ds, args):
    """ReSTIR GI temporal reuse (the reference's indirect method 2: one pass per frame tracing up to
    15 bounces from the primary vertex, merging into the reprojected previous reservoir, M-cap 40)
    at the bench resolution on the same terrain and camera walk; HIP-event times of its kernel."""
    import numpy as np
    import torch

    W, H = args.width, args.height
    L, H_ = rt.core(), rt.host()
    r = C.c_void_p()
    rt.check(H_.rt_renderer_create(ds.handle, W, H, C.byref(r)), "rt_renderer_create")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    gi = [torch.zeros(L.rt_gi_reservoir_bytes(W, H), dtype=torch.uint8, device="cuda") for _ in range(2)]
    for b in gi:
        rt.check(L.rt_gi_reservoir_reset(b.data_ptr(), W, H, stream), "reset")
    cur = [0]

    def frame(k, ev=None):
        cams = hs.camera(k % 16, W, H)
        f = rt.Frame()
        rt.check(H_.rt_renderer_bindings(r, C.byref(f)), "bindings")
        g = rt.GIReservoirs()
        g.reservoir, g.reservoirPrev = gi[cur[0]].data_ptr(), gi[cur[0] ^ 1].data_ptr()
        cur[0] ^= 1
        rt.check(L.rt_gbuffer_trace(ds.handle, C.byref(cams[0]), f.depthNormal, f.albedoMatId, f.motionVector,
                                    stream), "gbuffer")
        if ev: ev[0].record()
        rt.check(L.rt_gi_temporal(ds.handle, cams, C.byref(f), C.byref(g), stream), "gi")
        if ev: ev[1].record()
        rt.check(H_.rt_renderer_flip(r), "flip")

    try:
        for k in range(args.warmup):
            frame(k)
        torch.cuda.synchronize()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            frame(args.warmup + k, ev[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gi_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    finally:
        H_.rt_renderer_destroy(r)
    return {"ms_per_frame": round(wall * 1e3 / args.steps, 3), "gi_temporal_ms": round(gi_ms, 3),
            "mpixels_per_s": round(W * H * args.steps / wall / 1e6, 2),
            "workload": f"ReSTIR GI temporal (G-buffer + gi_resample_temporal, <= 15 bounces, M-cap 40), {W}x{H}, "
                        f"same terrain and camera walk"}


def progress(msg):
    """One line per finished step on stderr (the JSON result goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_sources_digest():
    """SHA-256 over the sources the kernels are built from (csrc/*.hip, csrc/*.h and
    include/restirpt.h, sorted by name): the build a traffic measurement belongs to. The host
    layer's header and sources are not part of it."""
    import hashlib

    h = hashlib.sha256()
    files = []
    csrc = os.path.join(ROOT, "vulkan-restir-pt_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".hip", ".h")):
            files.append((name, os.path.join(csrc, name)))
    files.append(("restirpt.h", os.path.join(ROOT, "include", "restirpt.h")))
    for name, path in files:
        h.update(name.encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def read_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, with the kernel-source digest
    it was measured at: (bytes, None) when the digest matches this tree, (None, reason) when the
    summary is missing or was measured on other kernel sources (a stale figure is not reported)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_1080p_terrain_traffic.md")
    if not os.path.exists(path):
        return None, "no traffic summary"
    val, digest = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                val = float(line.split(":", 1)[1].strip())
            except ValueError:
                val = None
        elif line.startswith("kernel_sources_sha256:"):
            digest = line.split(":", 1)[1].strip()
    if digest != kernel_sources_digest():
        return None, "the traffic summary was measured on other kernel sources (stale)"
    return val, None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs (--gpus > 1) are not part of it"}))
        sys.exit(2)
    import numpy as np
    import torch

    import restirpt as rt

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        copy_calibration()
        return
    if args.only_4k:
        print(json.dumps(full_form_4k(rt, args)))
        return
    if args.only_lbvh:
        print(json.dumps(lbvh_build(rt, [("terrain_1080p_scene", rt.HostScene(1, args.seed)),
                                         ("full_form_4k_scene", rt.HostScene(5, 4))])))
        return
    if args.only_plain:
        print(json.dumps(pt_plain_c0(rt, args)))
        return
    W, H = args.width, args.height
    hs = rt.HostScene(1, args.seed)
    ds = rt.DeviceScene(hs.arrays)
    info = ds.info()
    s = settings(rt, W)
    if args.learned_kernel_loop:
        print(json.dumps(learned_kernel_loop(rt, hs, ds, args)))
        return
    if args.di_kernel_loop:
        print(json.dumps(restir_di(rt, hs, ds, args, kernel_loop=True)))
        return
    if args.guided_kernel_loop:
        print(json.dumps(guided_kernel_loop(rt, hs, ds, args)))
        return
    if args.learned_study:
        print(json.dumps(learned_study(rt, hs, ds, args)))
        return
    if args.only_overlapped_loop:
        ov, _ = overlapped_frames(rt, hs, ds, W, H, s, args)
        print(json.dumps({"metric": METRIC, "value": ov["value"], "unit": "Mpaths/s", "ms_per_step": ov["ms_per_step"],
                          "frames_overlapped": ov}))
        return
    loop = FrameLoop(rt, hs, ds, W, H, s)
    for k in range(args.warmup):
        loop.frame(k)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        loop.frame(args.warmup + k, None if args.only_kernel_loop else ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    seq_out = np.zeros((H, W, 4), np.float32)
    rt.check(rt.host().rt_renderer_read_output(loop.r, 1, seq_out.ctypes.data, loop.stream), "read output")
    loop.close()
    ms_per_step = wall * 1e3 / args.steps
    mpaths = W * H * args.steps / wall / 1e6
    line = {"metric": METRIC, "value": round(mpaths, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "ReSTIR PT (GRIS) plain form, hybrid shift, temporal + spatial reuse "
                                   "(3 neighbours, 30 px, M-cap 20), 1920x1080, 1 path/px, max depth 6, procedural "
                                   "terrain (512x512-cell fBm heightfield, 4 tris/cell) + 200 instanced icospheres, "
                                   "16 emissive triangles, seeded 16-frame camera walk",
                       "width": W, "height": H, "triangles": int(info.numPrimitives), "emissive": 16,
                       "max_depth": 6, "bvh_depth": int(info.bvhDepth), "frames": "one after the other"}}
    if args.only_kernel_loop:
        print(json.dumps(line))
        return
    if not args.no_overlap:
        ov, ov_out = overlapped_frames(rt, hs, ds, W, H, s, args)
        ov["identical_to_sequential"] = bool(((ov_out == seq_out) | (np.isnan(ov_out) & np.isnan(seq_out))).all())
        ov["sequential_ms_per_step"] = line["ms_per_step"]
        ov["sequential_value"] = line["value"]
        line["frames_overlapped"] = ov
        # the frames are the same frames, bit for bit: the faster order is the throughput
        if ov["identical_to_sequential"] and ov["value"] > line["value"]:
            line["value"], line["ms_per_step"] = ov["value"], ov["ms_per_step"]
            line["config"]["frames"] = ("overlapped on two streams: frame k+1's G-buffer and path trace beside "
                                        "frame k's temporal and spatial reuse (rt_pipeline)")
        progress(f"overlapped frames: {ov['ms_per_step']} ms/frame (identical {ov['identical_to_sequential']})")
    pass_ms = {p: float(np.mean([ev[k][i].elapsed_time(ev[k][i + 1]) for k in range(args.steps)]))
               for i, p in enumerate(PASSES)}
    line["passes_ms"] = {p: round(v, 3) for p, v in pass_ms.items()}
    if not args.no_counters:
        work = count_work(rt, hs, ds, W, H, s, args.warmup, args.steps)
        rays = sum(int(work[p][2]) for p in PASSES)
        line["mrays_per_s"] = round(rays / wall / 1e6, 1)
        dom = max(PASSES, key=lambda p: pass_ms[p])
        nodes, tris = int(work[dom][0]), int(work[dom][1])
        per_launch = (IO_BYTES[dom] * W * H * args.steps + NODE_BYTES * nodes + TRI_BYTES * tris) / args.steps
        achieved = per_launch / (pass_ms[dom] * 1e-3) / 1e9
        line["roofline"] = {"bound": "hbm", "kernel": KERNELS[dom], "achieved": round(achieved, 1),
                            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                            "traffic": read_traffic(KERNELS[dom])[0],
                            "algorithmic_bytes_per_launch": int(per_launch),
                            "launch_ms": round(pass_ms[dom], 3)}
        stale = read_traffic(KERNELS[dom])[1]
        if stale:
            line["roofline"]["traffic_note"] = stale
        line["work_per_frame"] = {p: {"rays": int(work[p][2]) // args.steps, "nodes": int(work[p][0]) // args.steps,
                                      "tris": int(work[p][1]) // args.steps} for p in PASSES}
    progress(f"headline: {line['ms_per_step']} ms/frame")
    if not args.no_lbvh:
        line["lbvh_build"] = lbvh_build(rt, [("terrain_1080p_scene", hs), ("full_form_4k_scene", rt.HostScene(5, 4))])
        progress("lbvh_build done")
    if not args.no_plain:
        line["pt_plain_256"] = pt_plain_c0(rt, args)
        progress("pt_plain_256 done")
    if not args.no_di:
        line["restir_di"] = restir_di(rt, hs, ds, args)
        progress("restir_di done")
        line["restir_gi"] = restir_gi(rt, hs, ds, args)
        progress("restir_gi done")
        line["full_image"] = full_image(rt, hs, ds, args)
        progress("full_image done")
    if not args.no_batched:
        line["batched"] = batched_views(rt, args, hs, ds)
        progress("batched done")
    if not args.no_4k:
        line["full_form_4k"] = full_form_4k(rt, args)
        progress("full_form_4k done")
    if not args.no_learned:
        line["learned_vs_plain"] = learned_vs_plain(rt, hs, ds, args)
        progress("learned_vs_plain done")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(rt, hs, args)
        progress("cpu_baseline done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
