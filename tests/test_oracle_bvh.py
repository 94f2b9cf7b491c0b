"""Pins the oracle's LBVH and traversal: structural invariants, an independent
Morton-code computation, brute force closest hit, double-precision intersection
and watertightness."""
import numpy as np
import pytest

import orc
import rtpkg

rt = rtpkg.rt
LEAF = 0x80000000


@pytest.fixture(scope="module")
def boxroom():
    hs = rt.HostScene(0, 0)
    return hs, orc.OracleScene(hs.arrays)


def spread_bits(v):
    """Interleave helper written bit by bit (independent of the magic-number expansion)."""
    out = np.zeros_like(v, dtype=np.uint64)
    for b in range(21):
        out |= ((v.astype(np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def test_morton_codes_independent(boxroom):
    hs, osc = boxroom
    codes, order, nodes, box = osc.bvh()
    pos, ref = osc.world_tris()
    v = pos.reshape(-1, 3, 3)
    c = ((v[:, 0] + v[:, 1]) + v[:, 2]) * np.float32(1.0 / 3.0)
    lo, hi = c.min(0), c.max(0)
    assert np.array_equal(lo, box[:3]) and np.array_equal(hi, box[3:])
    ext = hi - lo
    scale = np.where(ext > 0, np.float32(2097151.0) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    q = np.minimum(((c - lo) * scale).astype(np.float32).astype(np.uint32), 2097151)
    mort = (spread_bits(q[:, 0]) << np.uint64(2)) | (spread_bits(q[:, 1]) << np.uint64(1)) | spread_bits(q[:, 2])
    # stable sort by code == lexicographic (code, primitive)
    want_order = np.lexsort((np.arange(len(mort)), mort))
    assert np.array_equal(order, want_order.astype(np.uint32))
    assert np.array_equal(codes, mort[want_order])


def test_lbvh_structure(boxroom):
    hs, osc = boxroom
    codes, order, nodes, box = osc.bvh()
    n = len(order)
    assert sorted(order.tolist()) == list(range(n))
    pos, _ = osc.world_tris()
    tri_lo = pos.reshape(-1, 3, 3).min(1)
    tri_hi = pos.reshape(-1, 3, 3).max(1)
    f = nodes[:, :12].view(np.float32)
    child = nodes[:, 12:14]
    seen_leaf = np.zeros(n, int)
    seen_node = np.zeros(n - 1, int)
    seen_node[0] = 1

    def subtree_box(ref):
        if ref & LEAF:
            i = int(ref) & 0x7FFFFFFF
            p = order[i]
            return tri_lo[p], tri_hi[p]
        return np.minimum(f[ref, 0:3], f[ref, 6:9]), np.maximum(f[ref, 3:6], f[ref, 9:12])

    for i in range(n - 1):
        for k, ref in enumerate(child[i]):
            if ref & LEAF:
                seen_leaf[int(ref) & 0x7FFFFFFF] += 1
            else:
                seen_node[ref] += 1
            lo, hi = subtree_box(int(ref))
            # the stored child box is exactly the union of the child's contents
            assert np.array_equal(f[i, 6 * k:6 * k + 3], lo) and np.array_equal(f[i, 6 * k + 3:6 * k + 6], hi)
    assert (seen_leaf == 1).all(), "every leaf must be referenced exactly once"
    assert (seen_node == 1).all(), "every internal node has exactly one parent"


def test_bvh4_collapse_structure(boxroom):
    """The traversed 4-wide tree covers every leaf once, its child boxes are the exact
    unions of their contents, and its closest hits equal the binary tree's."""
    hs, osc = boxroom
    codes, order, nodes, box = osc.bvh()
    n4 = osc.bvh4()
    assert 0 < len(n4) < len(nodes)
    pos, _ = osc.world_tris()
    tri_lo = pos.reshape(-1, 3, 3).min(1)
    tri_hi = pos.reshape(-1, 3, 3).max(1)
    f = n4[:, :24].view(np.float32).reshape(-1, 2, 3, 4)  # [node][lo/hi][axis][slot]
    child = n4[:, 24:28]
    leaves = np.zeros(len(order), int)
    parents = np.zeros(len(n4), int)
    parents[0] = 1

    def contents(ref):
        if int(ref) & LEAF:
            p = order[int(ref) & 0x7FFFFFFF]
            return tri_lo[p], tri_hi[p]
        lo = np.min(np.where(child[ref] != 0xFFFFFFFF, f[ref, 0], np.inf), axis=1)
        hi = np.max(np.where(child[ref] != 0xFFFFFFFF, f[ref, 1], -np.inf), axis=1)
        return lo, hi

    for i in range(len(n4)):
        for k in range(4):
            ref = child[i, k]
            if ref == 0xFFFFFFFF:
                assert np.isinf(f[i, :, :, k]).all()
                continue
            if int(ref) & LEAF:
                leaves[int(ref) & 0x7FFFFFFF] += 1
            else:
                parents[ref] += 1
            lo, hi = contents(ref)
            assert np.array_equal(f[i, 0, :, k], lo) and np.array_equal(f[i, 1, :, k], hi)
    assert (leaves == 1).all() and (parents == 1).all()
    g = np.random.default_rng(9)
    rays = np.zeros((20000, 8), np.float32)
    rays[:, :3] = g.uniform([-0.9, -0.9, 0.1], [0.9, 0.9, 1.9], (20000, 3))
    d = g.normal(size=(20000, 3))
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6], rays[:, 7] = 1e-4, 1e7
    t4, _, i4, _, s4 = osc.trace(rays, 0)
    t2, _, i2, _, s2 = osc.trace(rays, 3)
    assert np.array_equal(t4, t2)
    assert s4[1] < 0.75 * s2[1], "the 4-wide tree should visit far fewer nodes"


def test_closest_hit_matches_brute_force(boxroom):
    hs, osc = boxroom
    g = np.random.default_rng(5)
    n = 20000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = g.uniform([-0.9, -0.9, 0.1], [0.9, 0.9, 1.9], (n, 3))
    d = g.normal(size=(n, 3))
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6], rays[:, 7] = 1e-4, 1e7
    t, bary, inst, prim, st = osc.trace(rays, 0)
    tb, baryb, instb, primb, _ = osc.trace(rays, 2)
    assert np.array_equal(t, tb), "closest distance must equal brute force"
    diff = (inst != instb) | (prim != primb)
    # hit ids agree except exact-distance ties (counted)
    assert diff.sum() <= 5, f"{int(diff.sum())} ids differ at equal distance"
    assert (inst != 0xFFFFFFFF).mean() > 0.999  # the room is closed
    occ, *_ = osc.trace(np.concatenate([rays[:, :7], t[:, None] * 0.999], 1), 1)
    # any-hit with tmax just short of the closest hit finds nothing; beyond it finds something
    assert osc.trace(np.concatenate([rays[:, :7], (t * 0.999)[:, None]], 1), 1)[2].sum() <= 20
    assert osc.trace(np.concatenate([rays[:, :7], (t * 1.001 + 1e-4)[:, None]], 1), 1)[2].mean() > 0.999


def test_intersection_matches_double_precision(boxroom):
    hs, osc = boxroom
    pos, ref = osc.world_tris()
    g = np.random.default_rng(6)
    k = g.integers(0, len(pos), 3000)
    v = pos[k].reshape(-1, 3, 3).astype(np.float64)
    # rays from random origins through random interior points of the chosen triangle
    b = g.dirichlet([1, 1, 1], len(k))
    target = (b[:, :, None] * v).sum(1)
    o = target + g.normal(size=(len(k), 3)) * 0.3
    d = target - o
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    rays = np.zeros((len(k), 8), np.float32)
    rays[:, :3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, 0.0, 1e7
    t, bary, inst, prim, _ = osc.trace(rays, 2)
    hit_self = (inst == ref[k, 0]) & (prim == ref[k, 1])
    # when the target triangle is the closest hit, t and the barycentrics match double precision
    m = hit_self
    assert m.sum() > 500  # the others are occluded by nearer geometry
    # Moller-Trumbore in float64 on the float32-rounded ray and vertices
    O, D = rays[:, :3].astype(np.float64), rays[:, 3:6].astype(np.float64)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    pv = np.cross(D, e2)
    det = (e1 * pv).sum(1)
    tv = O - v[:, 0]
    u = (tv * pv).sum(1) / det
    qv = np.cross(tv, e1)
    w = (D * qv).sum(1) / det
    tt = (e2 * qv).sum(1) / det
    # float32 error of the sheared edge functions grows as |o| eps / |cos(incidence)|
    nrm = np.cross(e1, e2)
    cos = np.abs((nrm * D).sum(1)) / np.linalg.norm(nrm, axis=1)
    tol = 8 * np.abs(O).max(1) * 2.0 ** -24 / cos + 1e-7
    assert (np.abs(t - tt) <= tol)[m].all()
    well = m & (cos > 0.1)
    assert np.allclose(bary[well, 0], u[well], atol=2e-5) and np.allclose(bary[well, 1], w[well], atol=2e-5)


def test_watertight_shared_edges():
    """Rays aimed at the interior vertices (shared by 4-8 triangles) and edges (shared by
    two) of the terrain mesh never slip through."""
    hs = rt.HostScene(1, 1)
    osc = orc.OracleScene(hs.arrays)
    verts = hs.numpy()["vertices"]
    P = np.array([list(v) for v in verts["pos"][: 513 * 513 + 512 * 512]], np.float32)
    g = np.random.default_rng(7)
    i = g.integers(1, 511, 3000)
    j = g.integers(1, 511, 3000)
    corner = P[j * 513 + i]
    centre = P[513 * 513 + j * 512 + i]
    s = g.random((3000, 1)).astype(np.float32)
    edge = corner * (1 - s) + centre * s
    pts = np.concatenate([corner, centre, edge]).astype(np.float32)
    o = pts + np.array([0.013, -0.021, 3.0], np.float32)
    d = pts - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(pts), 8), np.float32)
    rays[:, :3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, 0.0, 1e7
    t, bary, inst, prim, _ = osc.trace(rays, 0, skip_lights=True)
    assert (inst != 0xFFFFFFFF).all(), f"{int((inst == 0xFFFFFFFF).sum())} rays slipped through shared edges"


def test_nonfinite_rays_miss_without_traversal(boxroom):
    """A ray with a NaN or infinite origin or direction hits nothing and visits no node (the
    slab test would otherwise accept every box for it and walk the whole tree)."""
    hs, osc = boxroom
    rays = np.zeros((6, 8), np.float32)
    rays[:, :3] = [0.1, 0.2, 0.5]
    rays[:, 3:6] = [0.0, 0.0, 1.0]
    rays[:, 6], rays[:, 7] = 1e-4, 1e7
    rays[0, 3] = np.nan
    rays[1, 0] = np.nan
    rays[2, 4] = np.inf
    rays[3, 2] = -np.inf
    rays[4, 3:6] = np.nan
    for mode in (0, 1, 3, 4):
        t, bary, inst, prim, st = osc.trace(rays[:5], mode)
        assert (inst[:5] == (0 if mode in (1, 4) else 0xFFFFFFFF)).all()
        assert st[1] == 0 and st[2] == 0 and st[0] == 5
    # the finite ray still hits the room
    _, _, inst, _, st = osc.trace(rays[5:], 0)
    assert inst[0] != 0xFFFFFFFF and st[1] > 0
