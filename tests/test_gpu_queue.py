"""Eager persistent path traces on one scene from two streams at once, more of them than the scene
has pixel-queue slots (64): each launch takes the next slot and first waits, on its own stream,
for the slot's previous launch, so no two launches in flight share a queue. Every launch renders
the same view into its own reservoir buffer; all of them must equal the first, bit for bit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402

rt = gpuhelp.rt


def test_concurrent_path_traces_never_share_a_queue():
    W, H = 96, 64
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    s = rt.settings(shift=2, max_depth=6)
    L = rt.core()
    rb = L.rt_gris_reservoir_bytes(W, H)
    launches = 2 * 64 + 6
    outs = [gpuhelp.dev_empty(rb) for _ in range(launches)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    print(f"[gpu] {launches} path traces on two streams", flush=True)
    for k in range(launches):
        f = run.frame_struct()
        f.grisReservoir = outs[k].data_ptr()
        st = C.c_void_p(streams[k % 2].cuda_stream)
        rt.check(L.rt_gris_path_trace(ds.handle, cams, C.byref(f), C.byref(s), st), "path trace")
    torch.cuda.synchronize()
    first = run.reservoirs(outs[0])
    assert (first["F"] != 0).any()  # the view has lit samples
    for k in range(1, launches):
        assert not gpuhelp.compare_reservoirs(first, run.reservoirs(outs[k])).any(), f"launch {k} differs"


def test_path_trace_result_does_not_depend_on_workgroup_count():
    W, H = 96, 64
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    run = gpuhelp.GrisRun(ds, W, H)
    cams = hs.camera(0, W, H)
    run.gbuffer(cams[0])
    s = rt.settings(shift=2, max_depth=6)
    L = rt.core()
    rb = L.rt_gris_reservoir_bytes(W, H)
    outs = {}
    for groups in (0, 1, 7, 100, 1 << 20):
        out = gpuhelp.dev_empty(rb)
        f = run.frame_struct()
        f.grisReservoir = out.data_ptr()
        rt.check(L.rt_gris_path_trace_groups(ds.handle, cams, C.byref(f), C.byref(s), groups, gpuhelp.stream()),
                 f"path trace, {groups} workgroups")
        outs[groups] = run.reservoirs(out)
    assert (outs[0]["F"] != 0).any()
    for groups, r in outs.items():
        assert not gpuhelp.compare_reservoirs(outs[0], r).any(), f"{groups} workgroups differ"
