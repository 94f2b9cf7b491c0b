"""rt_bvh_build_noalloc (the allocation-free binary LBVH build) against the build that
rt_device_scene_create runs, read back with rt_bvh_export: codes, order and nodes bit-identical,
on two calls in a row with the second on a smaller scene (the static scratch keeps the first
call's contents, which the build must not depend on), then the first scene again."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt


def prim_tris(hs):
    """World-space triangles in primitive order as the scene build lays them out: 3 float4 per
    primitive, .w = instance custom index, primitive index, 0 (float bits)."""
    pos, ref = orc.OracleScene(hs.arrays).world_tris()
    n = len(pos)
    t = np.zeros((n, 3, 4), np.float32)
    t[:, :, :3] = pos.reshape(n, 3, 3)
    t[:, 0, 3] = ref[:, 0].view(np.float32)
    t[:, 1, 3] = ref[:, 1].view(np.float32)
    return t


def build_noalloc(tris):
    n = tris.shape[0]
    d_tris = gpuhelp.from_host(tris)
    codes = torch.zeros(n, dtype=torch.int64, device="cuda")
    order = torch.zeros(n, dtype=torch.int32, device="cuda")
    nodes = torch.zeros(max(1, n - 1) * 16, dtype=torch.float32, device="cuda")
    st = rt.core().rt_bvh_build_noalloc(d_tris.data_ptr(), n, codes.data_ptr(), order.data_ptr(), nodes.data_ptr(),
                                        gpuhelp.stream())
    rt.check(st, "rt_bvh_build_noalloc")
    torch.cuda.synchronize()
    return (codes.cpu().numpy().view(np.uint64), order.cpu().numpy().view(np.uint32),
            nodes.cpu().numpy().view(np.uint32).reshape(-1, 16))


@pytest.fixture(scope="module")
def scenes():
    out = []
    for cfg in (0, 2):   # box room (20.5k triangles), then the floor scene (a few)
        hs = rt.HostScene(cfg, 0)
        out.append((cfg, hs, rt.DeviceScene(hs.arrays)))
    return out


def check_same(hs, ds):
    codes, order, nodes, info = gpuhelp.gpu_bvh(ds)
    tris = prim_tris(hs)
    assert tris.shape[0] == info.numPrimitives
    c2, o2, n2 = build_noalloc(tris)
    assert np.array_equal(c2, codes), "Morton codes differ"
    assert np.array_equal(o2, order), "order differs"
    assert np.array_equal(n2[:, :14], nodes), "nodes differ"


def test_noalloc_build_matches_scene_build(scenes):
    for cfg, hs, ds in scenes:
        print(f"[gpu] noalloc LBVH, scene config {cfg}", flush=True)
        check_same(hs, ds)
    cfg, hs, ds = scenes[0]
    print(f"[gpu] noalloc LBVH, scene config {cfg} again", flush=True)
    check_same(hs, ds)


def test_noalloc_build_rejects_sizes_beyond_its_scratch():
    L = rt.core()
    x = torch.zeros(16, device="cuda")
    p = x.data_ptr()
    assert L.rt_bvh_build_noalloc(p, (1 << 17) + 1, p, p, p, gpuhelp.stream()) == 4
    assert b"131072" in L.rt_last_error()
    assert L.rt_bvh_build_noalloc(p, 0, p, p, p, gpuhelp.stream()) == 4
    assert L.rt_bvh_build_noalloc(None, 1, p, p, p, gpuhelp.stream()) == 1
