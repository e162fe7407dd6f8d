"""Small scenes with known answers for next-event estimation tests.

Scenes are dicts in the layout mrh_scene_desc takes (tris, prim_key,
material tables, camera), shared by the oracle (CPU) and GPU tests."""
import math

import numpy as np

from meturay_amd import native as N


def quad(cx, cy, z, hx, hy):
    """Two triangles of the rectangle [cx-hx, cx+hx] x [cy-hy, cy+hy] at height z."""
    p = [(cx - hx, cy - hy, z), (cx + hx, cy - hy, z), (cx + hx, cy + hy, z), (cx - hx, cy + hy, z)]
    return [sum((p[0], p[1], p[2]), ()), sum((p[0], p[2], p[3]), ())]


def square_lights_scene(lights, rho=0.8, floor_half=3.0, mirror=False, reflect=None):
    """Lambert floor (albedo rho) at z = 0 under square lights given as
    (height, half size, radiance); a narrow camera halfway up to the lowest
    light looks straight down at the floor's centre.  With mirror the floor
    is a smooth conductor (roughness 0) instead; with reflect = (albedo rgb,
    roughness) a "Reflect" mirror."""
    tris = quad(0.0, 0.0, 0.0, floor_half, floor_half)
    batch = 5 if reflect is not None else (2 if mirror else 1)
    keys = [(batch << 24) | 0] * 2
    em = []
    for k, (z, half, le) in enumerate(lights):
        tris += quad(0.0, 0.0, z, half, half)
        keys += [(4 << 24) | k] * 2
        em.append([le, le, le])
    zc = 0.5 * min(z for z, _, _ in lights)
    return {
        "tris": np.array(tris, np.float32),
        "prim_key": np.array(keys, np.uint32),
        "lambert": np.array([[rho, rho, rho]], np.float32),
        "metal": np.array([[0.0, 0.2, 0.4, 1.4, 3.5, 2.4, 1.8, 1.0, 1.0, 1.0]], np.float32),
        "dielectric": np.zeros((0, 4), np.float32),
        "emissive": np.array(em, np.float32),
        "reflect": np.array([list(reflect[0]) + [reflect[1]]] if reflect is not None else np.zeros((0, 4)),
                            np.float32),
        "camera": {"position": [0.0, 0.0, zc], "gaze": [0.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0],
                   "fov_x": 0.02, "near": 1e-3, "far": 100.0},
    }


def view_factor_square(half, height):
    """Form factor from a differential area to a parallel, centred square of
    half size `half` at distance `height`: four corner rectangles of the
    textbook differential-to-parallel-rectangle factor."""
    a = half / height
    s = math.sqrt(1.0 + a * a)
    corner = (a / s) * math.atan(a / s) * 2.0 / (2.0 * math.pi)
    return 4.0 * corner


def furnace(light_radiance=1.0):
    """Camera inside a closed emissive box looking at a Lambert icosphere
    (albedo 0.5); every wall has radiance 1 except the light patch."""
    sc = N.gen_cornell(3)
    keys = sc["prim_key"].copy()
    keys[:12] = (4 << 24) | 0
    keys[12:14] = (4 << 24) | 1
    keys[14:] = (1 << 24) | 0
    sc["prim_key"] = keys
    sc["emissive"] = np.array([[1, 1, 1], [light_radiance] * 3], np.float32)
    sc["lambert"] = np.array([[0.5, 0.5, 0.5]], np.float32)
    return sc


def lit_spheres(reflect=False):
    """4x4 sphere grid (Lambert, GGX conductors of several roughnesses,
    dielectrics) with spheres 0 and 5 turned into lights; with reflect,
    spheres 2 and 9 become "Reflect" mirrors (one smooth, one rough)."""
    sc = N.gen_sphere_grid(4, 8, 8, 5)
    keys = sc["prim_key"].copy()
    per = len(keys) // 16
    keys[0:per] = (4 << 24) | 0
    keys[5 * per:6 * per] = (4 << 24) | 1
    sc["emissive"] = np.array([[4.0, 4.0, 4.0], [1.0, 2.0, 3.0]], np.float32)
    if reflect:
        keys[2 * per:3 * per] = (5 << 24) | 0
        keys[9 * per:10 * per] = (5 << 24) | 1
        sc["reflect"] = np.array([[0.9, 0.7, 0.5, 0.0], [1.0, 1.0, 1.0, 0.3]], np.float32)
    sc["prim_key"] = keys
    return sc


def smooth_normals(tris):
    """Per-corner shading normals [n][9] of a triangle soup: the area-weighted
    sum of the face normals of all triangles sharing a vertex position
    (bit-equal floats), normalised; a corner whose sum vanishes keeps its
    face normal."""
    t = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    face = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]).astype(np.float64)   # 2 x area
    corners = t.reshape(-1, 3)
    _, inv = np.unique(corners.view(np.uint32).view([("", np.uint32)] * 3), return_inverse=True)
    inv = inv.reshape(-1)
    acc = np.zeros((inv.max() + 1, 3))
    np.add.at(acc, inv, np.repeat(face, 3, axis=0))
    n = acc[inv]
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    fn = np.repeat(face / np.linalg.norm(face, axis=1, keepdims=True), 3, axis=0)
    n = np.where(ln > 1e-12, n / np.maximum(ln, 1e-300), fn)
    return n.reshape(-1, 9).astype(np.float32)


def smooth_spheres():
    """The 8 x 8 displaced-sphere grid with smooth vertex normals."""
    sc = N.gen_sphere_grid(8, 12, 12, 5)
    sc["normals"] = smooth_normals(sc["tris"])
    return sc


def uv_sphere(c, r, stacks=16, slices=24):
    """Triangles [..][9] of a UV sphere and their radial corner normals."""
    def pt(st, sl):
        th, ph = math.pi * st / stacks, 2.0 * math.pi * sl / slices
        return np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
    tris, nrm = [], []
    for st in range(stacks):
        for sl in range(slices):
            a, b, d, e = pt(st, sl), pt(st + 1, sl), pt(st + 1, sl + 1), pt(st, sl + 1)
            quads = []
            if st != 0:
                quads.append((a, e, b))      # counter-clockwise seen from outside
            if st != stacks - 1:
                quads.append((b, e, d))
            for q in quads:
                tris.append(np.concatenate([np.asarray(c) + r * v for v in q]))
                nrm.append(np.concatenate(q))
    return np.array(tris, np.float32), np.array(nrm, np.float32)


def smooth_cornell():
    """The Cornell box with three smooth UV spheres: Lambert, rough metal
    and glass; walls keep their face normals."""
    sc = N.gen_cornell(2)
    t = sc["tris"].reshape(-1, 3, 3).astype(np.float64)
    f = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    tris, nrm, keys = [sc["tris"]], [np.repeat(f, 3, axis=0).reshape(-1, 9)], [sc["prim_key"]]
    for c, r, key in [((0.5, -0.65, -0.3), 0.3, (1 << 24) | 0), ((-0.45, -0.6, -0.1), 0.35, (2 << 24) | 0),
                      ((0.0, 0.15, -0.45), 0.3, (3 << 24) | 0)]:
        st, sn = uv_sphere(c, r)
        tris.append(st)
        nrm.append(sn)
        keys.append(np.full(len(st), key, np.uint32))
    sc["tris"] = np.concatenate(tris).astype(np.float32)
    sc["normals"] = np.concatenate(nrm).astype(np.float32)
    sc["prim_key"] = np.concatenate(keys).astype(np.uint32)
    sc["metal"] = np.array([[0.25, 0.2, 0.4, 1.4, 3.5, 2.4, 1.8, 1.0, 1.0, 1.0]], np.float32)
    sc["dielectric"] = np.array([[1.0, 1.0, 1.0, 1.5]], np.float32)
    return sc
