"""Two-level (instanced) test scenes in the layout mrh_scene_desc takes:
tris / prim_key hold the meshes' local triangles and scene["instancing"]
places them (mesh_first, xform = local-to-world row-major 3x4, inst_mesh,
optional inst_key).  Shared by the oracle (CPU) and GPU tests."""
import math

import numpy as np

from meturay_amd import native as N


def uv_sphere(stacks=12, slices=16, r=1.0):
    """Triangles of a UV sphere of radius r at the origin."""
    tris = []
    for i in range(stacks):
        t0, t1 = math.pi * i / stacks, math.pi * (i + 1) / stacks
        for j in range(slices):
            p0, p1 = 2 * math.pi * j / slices, 2 * math.pi * (j + 1) / slices
            v = lambda t, p: (r * math.sin(t) * math.cos(p), r * math.sin(t) * math.sin(p), r * math.cos(t))
            a, b, c, d = v(t0, p0), v(t1, p0), v(t1, p1), v(t0, p1)
            if i > 0:
                tris.append(a + b + d)
            if i < stacks - 1:
                tris.append(b + c + d)
    return np.array(tris, np.float32)


def trs(t=(0, 0, 0), r=(0, 0, 0), s=(1, 1, 1)):
    """Row-major 3x4 local-to-world matrix: translate . rotate(x, then y, then z) . scale."""
    cx, sx = math.cos(r[0]), math.sin(r[0])
    cy, sy = math.cos(r[1]), math.sin(r[1])
    cz, sz = math.cos(r[2]), math.sin(r[2])
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    a = rz @ ry @ rx @ np.diag(s)
    return np.concatenate([a, np.array(t, float).reshape(3, 1)], 1).astype(np.float32).reshape(12)


def instanced_room(n_spheres=6, seed=5):
    """The Cornell room (one mesh, identity instance) holding n_spheres
    instances of one local UV sphere under rotations, non-uniform scales and
    translations, with Lambert / conductor / dielectric / Reflect keys given
    per instance, and two instances of a small emissive quad mesh (its
    triangles carry no emissive key: the instances' key override makes them
    lights)."""
    room = N.gen_cornell(3)
    rng = np.random.default_rng(seed)
    sphere = uv_sphere(10, 14, 1.0)
    quad = np.array([[-1, -1, 0, 1, -1, 0, 1, 1, 0], [-1, -1, 0, 1, 1, 0, -1, 1, 0]], np.float32)
    tris = np.concatenate([room["tris"], sphere, quad]).astype(np.float32)
    lam = np.concatenate([room["lambert"], np.array([[0.7, 0.6, 0.3]], np.float32)])
    keys = np.concatenate([room["prim_key"], np.full(len(sphere), N.key(1, len(lam) - 1), np.uint32),
                           np.full(len(quad), N.key(1, 0), np.uint32)]).astype(np.uint32)
    first = np.array([0, len(room["tris"]), len(room["tris"]) + len(sphere), len(tris)], np.uint32)
    xf = [trs()]
    mesh = [0]
    ikey = [0]
    kinds = [0, N.key(2, 0), N.key(3, 0), N.key(5, 0), N.key(1, 1), 0]
    for i in range(n_spheres):
        t = rng.uniform([-0.55, -0.75, -0.55], [0.55, -0.35, 0.35])
        s = rng.uniform(0.12, 0.25, 3)
        r = rng.uniform(-math.pi, math.pi, 3)
        xf.append(trs(t, r, s))
        mesh.append(1)
        ikey.append(kinds[i % len(kinds)])
    for i, x in enumerate((-0.4, 0.4)):
        xf.append(trs((x, 0.97, 0.1), (math.pi / 2, 0, 0.3 * i), (0.12, 0.08, 1.0)))
        mesh.append(2)
        ikey.append(N.key(4, i))
    return {
        "tris": tris,
        "prim_key": keys,
        "lambert": lam,
        "metal": np.array([[0.2, 0.2, 0.4, 1.4, 3.5, 2.4, 1.8, 1.0, 1.0, 1.0]], np.float32),
        "dielectric": np.array([[1.0, 1.0, 1.0, 1.5]], np.float32),
        "emissive": np.array([[6.0, 6.0, 5.0], [3.0, 4.0, 6.0]], np.float32),
        "reflect": np.array([[0.9, 0.9, 0.9, 0.0]], np.float32),
        "camera": room["camera"],
        "instancing": {"mesh_first": first, "xform": np.stack(xf).astype(np.float32),
                       "inst_mesh": np.array(mesh, np.uint32), "inst_key": np.array(ikey, np.uint32)},
    }


def split_identity(scene, cuts):
    """The single-level scene as a two-level one: meshes are the primitive
    ranges between cuts, each placed once by the identity."""
    n = len(scene["tris"])
    first = np.array([0] + list(cuts) + [n], np.uint32)
    out = dict(scene)
    out["instancing"] = {"mesh_first": first,
                         "xform": np.stack([trs()] * (len(first) - 1)).astype(np.float32),
                         "inst_mesh": np.arange(len(first) - 1, dtype=np.uint32)}
    return out


def flatten(scene):
    """World-space single-level scene of a two-level one (each instance's
    mesh triangles taken to the world in float64, keys overridden)."""
    inst = scene["instancing"]
    first, xf, im = inst["mesh_first"], inst["xform"].reshape(-1, 3, 4), inst["inst_mesh"]
    ik = inst.get("inst_key")
    tris, keys = [], []
    for i in range(len(im)):
        f, e = int(first[im[i]]), int(first[im[i] + 1])
        p = scene["tris"][f:e].reshape(-1, 3).astype(np.float64)
        w = p @ xf[i][:, :3].T.astype(np.float64) + xf[i][:, 3].astype(np.float64)
        tris.append(w.reshape(-1, 9).astype(np.float32))
        k = scene["prim_key"][f:e]
        if ik is not None and ik[i]:
            k = np.full(e - f, ik[i], np.uint32)
        keys.append(k)
    out = {k: v for k, v in scene.items() if k != "instancing"}
    out["tris"] = np.concatenate(tris)
    out["prim_key"] = np.concatenate(keys).astype(np.uint32)
    return out


def rays_into(scene, n, seed=3):
    """Rays from random points inside the room towards random directions
    (7 planes: o, d, tmax)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((7, n), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = 1e30
    return r
