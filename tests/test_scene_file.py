"""Scene files in the reference's XML format (Scene::load, src/Scene.cpp:104-303; materials,
src/Material.cpp:5-70; meshes as Resource::createNewModelInstance reads them through assimp,
src/Resource.cpp:100-262), loaded by rt_scene_load.

The fixtures under golden/scenes/ are written for these tests in that format (the reference ships
no scene files). What the loader must produce is pinned independently of its code:
* instance matrices against T * Rz(rx) * Rx(ry + 90) * Ry(rz) * S(sx, sz, sy) built from textbook
  rotation matrices (src/Model.cpp:11-21 composes glm::rotate calls in that order);
* light triangles against the OBJ corners pushed through that matrix by hand, the radiance against
  power / total area, normals against the flip attribute;
* smooth normals of the icosahedron against its radial directions (by symmetry the mean of the
  five faces around a vertex points away from the centre);
* the camera and transforms through the oracle's G-buffer: the left wall (red) fills the left edge
  of the image, the right wall (green) the right edge, the ball stands left of the box, the back
  wall lies at the distance the file puts it.
"""
import ctypes as C
import math
import shutil
from pathlib import Path

import numpy as np
import pytest

import rtpkg

rt = rtpkg.rt
SCENES = Path(rtpkg.ROOT) / "tests" / "golden" / "scenes"
ROOM = SCENES / "room.xml"


def rot(axis, deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = np.eye(4)
    if axis == "x":
        m[1:3, 1:3] = [[c, -s], [s, c]]
    elif axis == "y":
        m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    else:
        m[0:2, 0:2] = [[c, -s], [s, c]]
    return m


def model_matrix(t, s, r):
    T = np.eye(4)
    T[:3, 3] = t
    return T @ rot("z", r[0]) @ rot("x", r[1] + 90.0) @ rot("y", r[2]) @ np.diag([s[0], s[2], s[1], 1.0])


def obj_positions(path):
    return np.array([[float(x) for x in ln.split()[1:4]] for ln in open(path) if ln.startswith("v ")])


@pytest.fixture(scope="module")
def room():
    hs = rt.HostScene.load(ROOM)
    return hs, hs.numpy()


def test_counts_and_film(room):
    hs, d = room
    assert hs.film_size() == (96, 72)
    # room 5 quads, icosahedron 20 triangles, box 6 quads; lamp 1 quad
    assert [int(i["indexCount"]) // 3 for i in d["instances"]] == [10, 20, 12]
    assert len(d["lights"]) == 2 and len(d["lightTable"]) == 3
    # each face corner is a vertex of its own
    assert len(d["vertices"]) == 5 * 4 + 20 * 3 + 6 * 4
    assert len(d["materialIndices"]) == 42 and len(d["indices"]) == 126
    assert int(d["indices"].max()) < len(d["vertices"])


def test_materials(room):
    _, d = room
    m = d["materials"]
    base = np.array([tuple(x["baseColor"]) for x in m])
    # Resource's empty material, then per model its OBJ materials (assimp's DefaultMaterial first)
    assert np.allclose(base[0], [1, 0, 1]) and m[0]["type"] == 1
    assert np.allclose(base[1:5], [[0.6, 0.6, 0.6], [0.73, 0.73, 0.73], [0.63, 0.065, 0.05], [0.14, 0.45, 0.091]])
    assert list(m["type"][1:5]) == [1, 1, 1, 1]  # <material type="default"> keeps them
    # ball: metal with the XML's baseColor, roughness and ior
    assert m[5]["type"] == 3 and np.allclose(base[5], [0.9, 0.7, 0.3])
    assert np.isclose(m[5]["roughness"], 0.3) and np.isclose(m[5]["ior"], 2.5)
    # box: metalWorkflow without baseColor keeps the mesh material's colour
    assert m[6]["type"] == 2 and np.allclose(base[6], [0.6, 0.6, 0.6])
    assert np.isclose(m[6]["metallic"], 0.2) and np.isclose(m[6]["roughness"], 0.5) and np.isclose(m[6]["ior"], 1.5)
    assert (m["textureIdx"] == 0xFFFFFFFF).all()
    # per triangle: floor, ceiling, back wall white; left wall red; right wall green
    assert list(d["materialIndices"][:10]) == [2] * 6 + [3] * 2 + [4] * 2
    assert (d["materialIndices"][10:30] == 5).all() and (d["materialIndices"][30:] == 6).all()


@pytest.mark.parametrize("k,t,s,r", [(0, (0, 0, 0), (1, 1, 1), (0, 0, 0)),
                                     (1, (-0.4, 0.1, 0.35), (0.35, 0.35, 0.35), (0, 0, 0)),
                                     (2, (0.4, -0.2, 0), (0.3, 0.3, 0.6), (25, 0, 0))])
def test_instance_matrices(room, k, t, s, r):
    _, d = room
    inst = d["instances"][k]
    M = np.asarray(inst["transform"], np.float64).reshape(4, 4).T  # column-major
    assert np.allclose(M, model_matrix(t, s, r), atol=1e-6)
    Minv = np.asarray(inst["transformInv"], np.float64).reshape(4, 4).T
    assert np.allclose(Minv @ M, np.eye(4), atol=1e-5)
    assert np.allclose(np.asarray(inst["transformInvT"], np.float64).reshape(4, 4).T, Minv.T)


def test_rotate_attribute_order():
    """rotate="a b c": a about world z (after the y-up to z-up turn), b about x, c about the mesh's y."""
    M = model_matrix((0, 0, 0), (1, 1, 1), (90, 0, 0))
    # the mesh's +x axis turns to world +y under a 90 degree yaw
    assert np.allclose(M @ [1, 0, 0, 0], [0, 1, 0, 0])
    # the mesh's +y (up) axis is world +z
    assert np.allclose(model_matrix((0, 0, 0), (1, 1, 1), (0, 0, 0)) @ [0, 1, 0, 0], [0, 0, 1, 0])
    # scale "sx sy sz" scales world x, y, z
    assert np.allclose(model_matrix((0, 0, 0), (2, 3, 5), (0, 0, 0)) @ [1, 1, 1, 0], [2, -3, 5, 0])


def test_lights(room):
    _, d = room
    L = d["lights"]
    P = obj_positions(SCENES / "lamp.obj")
    M = model_matrix((0, 0, 0), (1, 1, 1), (0, 0, 0))
    W = (M @ np.c_[P, np.ones(4)].T).T[:, :3]
    tris = [(0, 3, 2), (0, 2, 1)]  # fan of "f 1 4 3 2"
    for l, (a, b, c) in zip(L, tris):
        assert np.allclose([l["v0"], l["v1"], l["v2"]], [W[a], W[b], W[c]], atol=1e-6)
        assert np.isclose(l["area"], 0.18, rtol=1e-5)
        # geometric normal +y in the mesh = world +z, flipped by flip="true"
        assert np.allclose([l["nx"], l["ny"], l["nz"]], [0, 0, -1], atol=1e-6)
        # power (8, 8, 6) spread over the lamp's 0.36 area
        assert np.allclose(l["radiance"], np.array([8, 8, 6]) / 0.36, rtol=1e-5)
    lum = 0.299 * 8 + 0.587 * 8 + 0.114 * 6
    assert np.isclose(d["lightTable"][0]["prob"], lum, rtol=1e-5) and d["lightTable"][0]["failId"] == 2


def test_smooth_normals_and_uvs(room):
    _, d = room
    v = d["vertices"]
    ball = v[20:80]  # the icosahedron's corners (its OBJ has no normals)
    pos = np.array([tuple(x["pos"]) for x in ball])
    nrm = np.array([tuple(x["norm"]) for x in ball])
    assert np.allclose(np.linalg.norm(pos, axis=1), 1.0, atol=1e-6)
    assert np.allclose(nrm, pos, atol=1e-6)
    # floor corners "1/1/1 4/4/1 3/3/1 2/2/1": UVs flipped (v -> 1 - v), given normal +y
    floor = v[:4]
    assert np.allclose(floor["uvx"], [0, 0, 1, 1]) and np.allclose(floor["uvy"], [1, 0, 0, 1])
    assert np.allclose([tuple(x["norm"]) for x in floor], [[0, 1, 0]] * 4)
    # corners without a vt read uv 0
    assert (v[4:20]["uvx"] == 0).all() and (v[4:20]["uvy"] == 0).all()


def test_camera(room):
    hs, _ = room
    a, b = hs.camera(0, 96, 72), hs.camera(5, 96, 72)
    c = a[0]
    assert np.allclose(list(c.pos), [0, -2.6, 1]) and np.allclose(list(c.front), [0, 1, 0], atol=1e-6)
    assert np.allclose(list(c.up), [0, 0, 1], atol=1e-6) and c.FOV == 55.0
    assert c.lensRadius == 0.0 and c.focalDist == 1.0 and list(c.filmSize) == [96, 72]
    # a still view: the same matrices every frame, a new seed per frame
    assert list(b[0].projView) == list(c.projView) and list(b[0].lastProjView) == list(c.projView)
    assert b[0].seed != c.seed


def test_gbuffer_sees_the_room(room):
    """The oracle's G-buffer of the loaded scene: walls, floor and objects where the file puts them."""
    import orc

    hs, _ = room
    osc = orc.OracleScene(hs.arrays)
    cam = hs.camera(0, 96, 72)[0]
    dn, am, _, _ = osc.gbuffer(cam)
    mat = am[..., 1] >> 16
    mid = 36
    assert (mat[mid, :16] == 3).all(), "left edge is not the red wall"
    assert (mat[mid, -16:] == 4).all(), "right edge is not the green wall"
    assert (mat[-2, 40:52] == 2).all(), "bottom centre is not the white floor"
    assert (mat[0, 30:66] == 2).all() and np.allclose(dn[0, 30:66, 1:], [0, 0, -1], atol=1e-6), "top is not the ceiling"
    assert (mat == 5).sum() > 30 and (mat == 6).sum() > 30, "ball and box not visible"
    # ball on the left half, box on the right half
    assert np.nonzero(mat == 5)[1].mean() < 48 < np.nonzero(mat == 6)[1].mean()
    # the back wall (world y = 1, normal -y) lies 3.6 ahead of the camera at the image centre
    assert mat[mid, 48] == 2 and abs(dn[mid, 48, 0] - 3.6) < 1e-3
    assert np.allclose(dn[mid, 48, 1:], [0, -1, 0], atol=1e-6)


def write_scene(tmp_path, models, camera='<position value="0 -3 1"/><lookAt value="0 0 1"/><fov value="45"/>'):
    for f in ("cube.obj", "lamp.obj"):
        shutil.copy(SCENES / f, tmp_path / f)
    xml = (f'<scene><integrator><size width="32" height="16"/></integrator><camera>{camera}</camera>'
           f'<modelInstances>{models}</modelInstances></scene>')
    p = tmp_path / "s.xml"
    p.write_text(xml)
    return p


LAMP = '<model type="light" path="lamp.obj"><radiance value="1 1 1"/></model>'


def test_material_types(tmp_path):
    kinds = [('principled', 1, {}), ('lambertian', 1, {}), ('thinDielectric', 5, {"ior": 1.3}),
             ('dielectric', 4, {"ior": 1.7, "roughness": 0.0}), ('fake', 6, {}),
             ('metal', 3, {"roughness": 0.05, "ior": 3.0}), ('metalWorkflow', 2, {"metallic": 1.0})]
    models = ""
    for name, _, params in kinds:
        body = "".join(f'<{k} value="{v}"/>' for k, v in params.items())
        models += f'<model path="cube.obj"><material type="{name}"><baseColor value="0.5 0.25 1"/>{body}</material></model>'
    models += '<model path="cube.obj"><material type="default"><baseColor value="0 0 0"/></material></model>'
    hs = rt.HostScene.load(write_scene(tmp_path, models + LAMP))
    m = hs.numpy()["materials"]
    assert len(m) == 1 + len(kinds) + 1
    for k, (name, typ, params) in enumerate(kinds):
        x = m[1 + k]
        assert x["type"] == typ, name
        assert np.allclose(x["baseColor"], [0.5, 0.25, 1.0]), name
        for key, val in params.items():
            assert np.isclose(x[key], val), (name, key)
        for key, dflt in (("metallic", 0.0), ("roughness", 1.0), ("ior", 1.5)):
            if key not in params:
                assert np.isclose(x[key], dflt), (name, key)
    # "default" keeps the mesh's material (assimp's DefaultMaterial, grey 0.6)
    assert m[-1]["type"] == 1 and np.allclose(m[-1]["baseColor"], 0.6)


def test_transform_defaults_and_polygons(tmp_path):
    (tmp_path / "penta.obj").write_text("v 0 0 0\nv 1 0 0\nv 1.5 0 1\nv 0.5 0 2\nv -0.5 0 1\nf 1 2 3 4 5\n")
    shutil.copy(SCENES / "lamp.obj", tmp_path / "lamp.obj")
    # no transform node: translate 0, rotate 0, scale 1
    (tmp_path / "s.xml").write_text('<scene><camera><position value="0 -3 1"/><angle value="90 0 0"/></camera>'
                                    '<modelInstances><model path="penta.obj"/>' + LAMP + '</modelInstances></scene>')
    hs = rt.HostScene.load(tmp_path / "s.xml")
    d = hs.numpy()
    assert np.allclose(np.asarray(d["instances"][0]["transform"]).reshape(4, 4).T,
                       model_matrix((0, 0, 0), (1, 1, 1), (0, 0, 0)), atol=1e-7)
    # a pentagon is fanned into three triangles around its first corner
    assert list(d["indices"]) == [0, 1, 2, 0, 2, 3, 0, 3, 4]
    # the smooth normal of a flat polygon is its plane normal (+y in the mesh, facing the winding)
    n = np.array([tuple(x["norm"]) for x in d["vertices"]])
    assert np.allclose(np.abs(n), [[0, 1, 0]] * 5, atol=1e-6)
    # angle "yaw pitch roll": front = (sin yaw cos pitch, cos yaw cos pitch, sin pitch) (src/Camera.cpp:79-83)
    c = hs.camera(0, 8, 8)[0]
    assert np.allclose(list(c.front), [1, 0, 0], atol=1e-6)
    (tmp_path / "s.xml").write_text('<scene><camera><position value="0 -3 1"/><angle value="30 20 0"/></camera>'
                                    '<modelInstances><model path="penta.obj"/>' + LAMP + '</modelInstances></scene>')
    c = rt.HostScene.load(tmp_path / "s.xml").camera(0, 8, 8)[0]
    y, p = math.radians(30), math.radians(20)
    assert np.allclose(list(c.front), [math.sin(y) * math.cos(p), math.cos(y) * math.cos(p), math.sin(p)], atol=1e-6)


@pytest.mark.parametrize("case,text,needle", [
    ("missing file", None, "cannot open"),
    ("not a scene", "<notascene/>", "failed to load"),
    ("unterminated", "<scene><modelInstances></scene>", "does not close"),
    ("missing texture", '<scene><modelInstances><model path="cube.obj"><material type="lambertian">'
                        '<baseColor image="wood.png"/></material></model></modelInstances></scene>', "png: cannot open"),
    ("zero power", '<scene><modelInstances><model type="light" path="lamp.obj"><radiance value="0 0 0"/>'
                   '</model></modelInstances></scene>', "no power"),
    ("bad index", '<scene><modelInstances><model path="bad.obj"/></modelInstances></scene>', "out of range"),
    ("missing map_Kd", '<scene><modelInstances><model path="tex.obj"/></modelInstances></scene>', "png: cannot open"),
    ("missing mesh", '<scene><modelInstances><model path="nothere.obj"/></modelInstances></scene>', "cannot open"),
    ("camera", '<scene><camera><position value="1 1 1"/><lookAt value="1 1 1"/></camera></scene>', "lookAt"),
])
def test_errors(tmp_path, case, text, needle):
    for f in ("cube.obj", "lamp.obj"):
        shutil.copy(SCENES / f, tmp_path / f)
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    (tmp_path / "tex.obj").write_text("mtllib tex.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    (tmp_path / "tex.mtl").write_text("newmtl a\nKd 1 1 1\nmap_Kd a.png\n")
    p = tmp_path / "s.xml"
    if text is not None:
        p.write_text(text)
    with pytest.raises(RuntimeError, match=needle):
        rt.HostScene.load(p)


def test_loaded_scene_uploads(room):
    """The loaded arrays pass the device scene's validation (checked on the host before any upload)."""
    hs, d = room
    a = hs.arrays
    assert a.numMaterialIndices * 3 == a.numIndices
    assert (d["materialIndices"] < a.numMaterials).all() and (d["materialIndices"] >= 0).all()
    assert C.sizeof(rt.Material) == 32
