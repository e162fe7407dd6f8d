"""Scene files (mrh_scene_file_load): the reference's JSON scene format with
OBJ meshes, checked against an independent reading of the same fixtures.

The fixtures under tests/golden/scenes were written for these tests.  The
expected triangles are rebuilt here from the OBJ text and the JSON with
numpy (and scipy's Euler-angle rotations for the trs transforms), so a
wrong transform order, a wrong matrix major, a wrong mesh split or a wrong
material / primitive pairing changes the comparison.
"""
import json
import math
import os
import re

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from meturay_amd import native as N

HERE = os.path.join(os.path.dirname(__file__), "golden", "scenes")
REF = "/root/reference/WorkingDir"


def obj_meshes(path):
    """OBJ meshes (positions per triangle) split at o / g / usemtl."""
    verts, meshes, cur = [], [], []
    faces = []
    for line in open(path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "v":
            verts.append([float(x) for x in t[1:4]])
        elif t[0] in ("o", "g", "usemtl"):
            if cur:
                faces.append(cur)
            cur = []
        elif t[0] == "f":
            idx = [int(x.split("/")[0]) for x in t[1:]]
            idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
            for k in range(1, len(idx) - 1):
                cur.append((idx[0], idx[k], idx[k + 1]))
    if cur:
        faces.append(cur)
    v = np.array(verts)
    return [np.array([[v[a], v[b], v[c]] for a, b, c in f]).reshape(-1, 9) for f in faces]


def strip_comments(text):
    text = re.sub(r"//[^\n]*", "", text)
    return re.sub(r",(\s*[\]}])", r"\1", text)


def matrices(scene):
    out = {}
    for node in scene["Transforms"]:
        ids = node["id"] if isinstance(node["id"], list) else [node["id"]]
        for i, tid in enumerate(ids):
            pick = (lambda k: node[k][i]) if len(ids) > 1 else (lambda k: node[k])
            if node["type"] == "Identity":
                out[tid] = np.eye(4)
            elif node["layout"] == "trs":
                m = np.eye(4)
                m[:3, :3] = Rotation.from_euler("xyz", pick("rotate"), degrees=True).as_matrix() @ np.diag(
                    pick("scale"))
                m[:3, 3] = pick("translate")
                out[tid] = m
            else:
                m = np.array(pick("matrix"), float).reshape(4, 4)
                out[tid] = m if node.get("major", "row") == "row" else m.T
    return out


def expected_room():
    scene = json.loads(strip_comments(open(os.path.join(HERE, "room.json")).read()))
    meshes = obj_meshes(os.path.join(HERE, "room_parts.obj"))
    mats = matrices(scene)
    prims = {0: meshes[0], 1: meshes[1], 2: meshes[2], 6: meshes[3],
             3: np.array(scene["Primitives"][2]["position"], float).reshape(-1, 9)}
    pos = np.array(scene["Primitives"][3]["position"], float)
    for pid, idx in zip([4, 5], scene["Primitives"][3]["index"]):
        prims[pid] = pos[np.array(idx)].reshape(-1, 9)
    keys = {1: N.key(1, 0), 2: N.key(1, 1), 3: N.key(5, 0), 4: N.key(3, 0), 5: N.key(2, 0), 6: N.key(1, 2)}
    tris, kk = [], []

    def emit(tri9, m, key):
        p = tri9.reshape(-1, 3, 3)
        h = np.concatenate([p, np.ones(p.shape[:2] + (1,))], axis=2) @ m.T
        tris.append(h[..., :3].reshape(-1, 9))
        kk.extend([key] * len(p))

    for s in scene["Surfaces"]:
        ms = s["material"] if isinstance(s["material"], list) else [s["material"]]
        ps = s["primitive"] if isinstance(s["primitive"], list) else [s["primitive"]]
        for m, p in zip(ms, ps):
            emit(prims[p], mats[s["transform"]], keys[m])
    emit(prims[6], np.eye(4), N.key(4, 0))
    return np.concatenate(tris), np.array(kk, np.uint32), scene


def test_room_triangles_keys_and_tables():
    sc = N.load_scene_file(os.path.join(HERE, "room.json"))
    tris, keys, scene = expected_room()
    assert sc["warnings"] == ""
    assert np.array_equal(sc["prim_key"], keys)
    assert np.allclose(sc["tris"], tris, atol=2e-6)
    assert np.allclose(sc["lambert"], [[0.7, 0.7, 0.7], [0.6, 0.2, 0.1], [0.2, 0.5, 0.3]])
    assert np.allclose(sc["reflect"], [[0.9, 0.9, 0.8, 0.0]])
    assert np.allclose(sc["dielectric"], [[1, 1, 1, 1.45]])
    assert np.allclose(sc["metal"], [[0.3, 0.2, 0.4, 1.4, 3.5, 2.4, 1.8, 1, 1, 1]])
    assert np.allclose(sc["emissive"], [[12, 10, 8]])
    # cameras in CameraSurfaces order, fov in radians
    c3, c0 = sc["cameras"]
    assert np.allclose(c3["position"], [3, 3, 3]) and c3["fov_x"] == pytest.approx(math.radians(40))
    assert np.allclose(c0["gaze"], [0, 1, 0]) and (c0["near"], c0["far"]) == pytest.approx((0.01, 50.0))


def test_obj_mesh_split_and_fans():
    import ctypes as C
    n = C.c_uint32()
    t = (C.c_uint32 * 8)()
    assert N.lib().mrh_obj_mesh_counts(os.path.join(HERE, "room_parts.obj").encode(), C.byref(n), t, 8) == 0
    assert n.value == 4 and list(t)[:4] == [2, 12, 3, 2]
    assert [len(m) for m in obj_meshes(os.path.join(HERE, "room_parts.obj"))] == [2, 12, 3, 2]


def test_unsupported_features_strict_and_substituted():
    path = os.path.join(HERE, "room_extras.json")
    with pytest.raises(N.MrhError) as e:
        N.load_scene_file(path)
    assert e.value.code == 4 and "not supported" in str(e.value)
    sc = N.load_scene_file(path, substitute=True)
    w = sc["warnings"]
    for what in ["Point", "Spherical", "textured albedo", "Unreal", "analytic sphere"]:
        assert what in w, what
    # the sphere became 64 x 32 - 64 triangles of the Unreal's albedo, the floor a grey Lambert
    assert len(sc["tris"]) == 64 * 32 * 2 - 2 * 64 + 2
    assert np.allclose(sc["lambert"], [[0.5, 0.5, 0.5], [0.3, 0.6, 0.9]])
    r = np.linalg.norm(sc["tris"][:-2].reshape(-1, 3) - [0, 1, 0], axis=1)
    assert np.allclose(r, 0.5, atol=1e-6)
    assert len(sc["cameras"]) == 1 and len(sc["emissive"]) == 0


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_malformed_files(tmp_path):
    bad = write(tmp_path, "bad.json", '{"Cameras": [\n{"id": 0 "type": "Pinhole"}]}')
    with pytest.raises(N.MrhError) as e:
        N.load_scene_file(bad)
    assert e.value.code == 1 and "line 2" in str(e.value)
    base = open(os.path.join(HERE, "room.json")).read()
    for old, new, msg in [('"material": 1, "primitive": 0', '"material": 9, "primitive": 0', "unknown material"),
                          ('"room_parts.obj", "tag": "assimp_", "innerIndex": 3', '"missing.obj", "innerIndex": 0',
                           "cannot read OBJ"),
                          ('"innerIndex": 3', '"innerIndex": 7', "has no mesh"),
                          ('"medium": 1', '"medium": 4', "unknown medium")]:
        assert old in base
        p = write(tmp_path, "room.json", base.replace(old, new))
        (tmp_path / "room_parts.obj").write_text(open(os.path.join(HERE, "room_parts.obj")).read())
        with pytest.raises(N.MrhError) as e:
            N.load_scene_file(p)
        assert msg in str(e.value), str(e.value)
    with pytest.raises(N.MrhError):
        N.load_scene_file(str(tmp_path / "nothing.json"))


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's scene files are not on this machine")
def test_reference_scenes_load():
    box = N.load_scene_file(os.path.join(REF, "TestScenes", "helloBox.json"))
    # four meshes of 6 + 2 + 2 + 2 indexed triangles, four LambertC albedos
    assert len(box["tris"]) == 12 and len(box["lambert"]) == 4
    assert sorted(set((box["prim_key"] & 0xFFFFFF).tolist())) == [0, 1, 2, 3]
    door = N.load_scene_file(os.path.join(REF, "Scenes", "veachDoor.json"), substitute=True)
    # every Door OBJ mesh of the surfaces and the lamp, independently counted
    d = os.path.join(REF, "Surfaces", "Door")
    count = lambda f, k=None: (sum(len(m) for m in obj_meshes(os.path.join(d, f))) if k is None
                               else len(obj_meshes(os.path.join(d, f))[k]))
    want = (count("table.obj") + count("table_legs.obj") + count("teapots.obj") + count("door_frames.obj") +
            count("door.obj") + count("door_hinges.obj") + count("door_knob.obj") + count("floor.obj") +
            count("floor_back.obj") + count("picture_frame.obj") + count("picture_image.obj") + count("walls.obj") +
            count("lamp.obj"))
    assert len(door["tris"]) == want
    assert np.allclose(door["emissive"], [[1420, 1552, 1642]])
    assert len(door["cameras"]) == 3 and "Spherical" in door["warnings"]
    with pytest.raises(N.MrhError):
        N.load_scene_file(os.path.join(REF, "Scenes", "veachDoor.json"))


def test_room_instanced_mode():
    """MRH_SCENE_FILE_INSTANCED: each used primitive becomes one local-space
    mesh and each surface primitive an instance; placing the instances
    (float64 product) gives the world triangles and keys of the baked load."""
    import instancing_scenes as I
    sc = N.load_scene_file(os.path.join(HERE, "room.json"), instanced=True)
    tris, keys, _ = expected_room()
    inst = sc["instancing"]
    assert inst["mesh_first"][0] == 0 and inst["mesh_first"][-1] == len(sc["tris"])
    # primitives shared by several surfaces are stored once
    assert len(inst["inst_mesh"]) > len(inst["mesh_first"]) - 1
    assert len(sc["tris"]) < len(tris)
    flat = I.flatten(sc)
    assert np.array_equal(flat["prim_key"], keys)
    assert np.allclose(flat["tris"], tris, atol=2e-6)
    # one instance per surface primitive (8) and per light surface (1), in
    # file order; primitives 4 and 5 are used twice, so 7 meshes
    assert len(inst["inst_mesh"]) == 9 and len(inst["mesh_first"]) - 1 == 7
    assert inst["inst_mesh"][6] == inst["inst_mesh"][5] and inst["inst_mesh"][7] == inst["inst_mesh"][4]
    assert ((inst["inst_key"] >> 24) == 4).tolist() == [False] * 8 + [True]


SMOOTH_OBJ = """# corner normals, a corner without one, a degenerate face
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
vn 0.3 0 1
vn 0 0.3 2
vn -0.2 -0.2 1
o a
f 1//1 2//2 3//3
f 1/1/4 3//3 4
f 1//1 2//2 1//1
"""

SMOOTH_JSON = """{
  "Transforms": [
    {"id": 0, "type": "Identity"},
    {"id": 1, "type": "Single", "layout": "trs", "translate": [0.5, 0.0, -0.25],
     "rotate": [20.0, 30.0, 40.0], "scale": [0.5, 2.0, 0.6]},
    {"id": 2, "type": "Single", "layout": "matrix4x4", "major": "row",
     "matrix": [-1, 0, 0, 0,  0, 1, 0.5, 0,  0, 0, 3, 1,  0, 0, 0, 1]}
  ],
  "Materials": [{"id": 1, "type": "LambertC", "albedo": [0.5, 0.5, 0.5]}],
  "Primitives": [
    {"id": 0, "type": "Triangle", "name": "smooth.obj", "tag": "assimp_", "innerIndex": 0},
    {"id": 1, "type": "Triangle", "name": ".nodeTriangle",
     "position": [[0, 0, 1], [1, 0, 1], [0, 1, 1]], "normal": [[0, 0, 1], [0.1, 0, 1], [0, 0.1, 1]]},
    {"id": 2, "type": "Triangle", "name": ".nodeTriangleIndexed",
     "position": [[0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2]],
     "normal": [[0, 0, 1], [0, 0.2, 1], [0.2, 0, 1], [0.1, 0.1, 1]], "index": [0, 1, 2, 0, 2, 3]},
    {"id": 3, "type": "Triangle", "name": ".nodeTriangle",
     "position": [[0, 0, 3], [1, 0, 3], [0, 1, 3]]}
  ],
  "Surfaces": [
    {"transform": 0, "material": 1, "primitive": 3},
    {"transform": 1, "material": 1, "primitive": 0},
    {"transform": 2, "material": 1, "primitive": 1},
    {"transform": 2, "material": 1, "primitive": 2}
  ]
}"""


def test_vertex_normals(tmp_path):
    """Corner normals from OBJ vn and the JSON "normal" lists, moved to world
    space by the inverse transpose (independently: numpy's inverse) and
    renormalised; a corner without a normal, and a flat primitive emitted
    before the first smooth one, get the world face normal; the degenerate
    face is dropped together with its normals."""
    write(tmp_path, "smooth.obj", SMOOTH_OBJ)
    p = write(tmp_path, "smooth.json", SMOOTH_JSON)
    sc = N.load_scene_file(p)
    tris = sc["tris"].reshape(-1, 3, 3).astype(np.float64)
    nrm = sc["normals"].reshape(-1, 3, 3).astype(np.float64)
    assert len(tris) == 1 + 2 + 1 + 2 and nrm.shape == tris.shape
    unit = lambda v: v / np.linalg.norm(v)
    face = lambda t: unit(np.cross(t[1] - t[0], t[2] - t[0]))
    M1 = matrices(json.loads(strip_comments(SMOOTH_JSON)))[1][:3, :3]
    M2 = np.array([[-1, 0, 0], [0, 1, 0.5], [0, 0, 3]], np.float64)
    vn = np.array([[0, 0, 1], [0.3, 0, 1], [0, 0.3, 2], [-0.2, -0.2, 1]], np.float64)
    world = lambda M, n: unit(np.linalg.inv(M).T @ n)
    want = [
        [face(tris[0])] * 3,                                         # flat, before the first smooth one
        [world(M1, vn[0]), world(M1, vn[1]), world(M1, vn[2])],
        [world(M1, vn[3]), world(M1, vn[2]), face(tris[2])],        # v/t/n, v//n, a corner without vn
        [world(M2, n) for n in ([0, 0, 1], [0.1, 0, 1], [0, 0.1, 1])],
        [world(M2, n) for n in ([0, 0, 1], [0, 0.2, 1], [0.2, 0, 1])],
        [world(M2, n) for n in ([0, 0, 1], [0.2, 0, 1], [0.1, 0.1, 1])],
    ]
    assert np.allclose(nrm, np.array(want), atol=2e-6)
    assert np.allclose(np.linalg.norm(nrm, axis=2), 1.0, atol=1e-6)
    # the mirror keeps normals on the side the local normal points to
    assert np.all(np.einsum("ij,ij->i", nrm[3], [face(tris[3])] * 3) < 0)
    # instanced: flat shading, noted
    inst = N.load_scene_file(p, instanced=True)
    assert "normals" not in inst and "flat shading" in inst["warnings"]
    # a scene without any normals stays flat
    flat = json.loads(SMOOTH_JSON)
    flat["Surfaces"] = flat["Surfaces"][:1]
    plain = N.load_scene_file(write(tmp_path, "flat.json", json.dumps(flat)))
    assert "normals" not in plain and len(plain["tris"]) == 1
    # malformed: a normal list of the wrong length
    bad = write(tmp_path, "bad_n.json", SMOOTH_JSON.replace('[0, 0.1, 1]]}', '[0, 0.1, 1], [0, 0, 1]]}'))
    with pytest.raises(N.MrhError) as e:
        N.load_scene_file(bad)
    assert "one normal per position" in str(e.value)


def test_substituted_sphere_is_smooth():
    sc = N.load_scene_file(os.path.join(HERE, "room_extras.json"), substitute=True)
    p = sc["tris"].reshape(-1, 3)
    n = sc["normals"].reshape(-1, 3)
    k = len(p) - 6                                                   # the floor's two triangles last
    assert np.allclose(n[:k], (p[:k] - [0, 1, 0]) / 0.5, atol=2e-6)   # radial
    t = p[:k].reshape(-1, 3, 3)
    f = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])                # wound to face outwards
    assert np.all(np.einsum("ij,ij->i", f, n[:k].reshape(-1, 3, 3).mean(axis=1)) > 0)
    assert np.allclose(n[k:], [0, -1, 0])                            # the floor's vn


def test_room_normals():
    # room_parts.obj's vn: the floor (mesh 0) and the pentagon (mesh 2) carry
    # one each, every other triangle its world face normal
    sc = N.load_scene_file(os.path.join(HERE, "room.json"))
    tris, _, scene = expected_room()
    t = tris.reshape(-1, 3, 3)
    f = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    want = np.repeat(f[:, None, :], 3, axis=1)
    want[0:2] = [0, -1, 0]                                           # floor, identity transform
    m1 = np.linalg.inv(matrices(scene)[1][:3, :3]).T @ [0, 0, 1]     # pentagon, transform 1
    want[14:17] = m1 / np.linalg.norm(m1)
    assert np.allclose(sc["normals"].reshape(-1, 3, 3), want, atol=2e-6)
