"""Two-level scenes in the oracle, pinned by what does not depend on it:

* the ray transform against the reference's own Ray::Transform (host build
  of RayLib, when present) bit for bit;
* the inverse by its product with the matrix, and exactly for translations
  and power-of-two scales;
* the instance boxes by containing every transformed vertex and touching
  each face;
* the traversal by brute force in each instance's local space (every
  triangle of every instance, the same local rays): the closest t is
  bit-identical, so are the primitive, instance and barycentrics;
* identity instances of a split single-level scene against the
  single-level tree: hits and renders bit-identical (the light list comes
  out in the same order), and with one mesh the node counts equal too.
"""
import numpy as np
import pytest

import instancing_scenes as I
from meturay_amd import native as N


def test_ray_transform_matches_reference(oracle):
    if not oracle.ref_available():
        pytest.skip("reference host build not present")
    R = oracle.ref()
    rng = np.random.default_rng(11)
    for k in range(400):
        m = rng.normal(size=12).astype(np.float32)
        if k % 7 == 0:
            m[[1, 2, 4, 6, 8, 9]] = 0.0      # axis-aligned: zero terms in the sums
        o = rng.normal(size=3).astype(np.float32) * 10
        d = rng.normal(size=3).astype(np.float32)
        if k % 5 == 0:
            o[0], d[1] = -0.0, -0.0
        ro, rd = oracle.xf_ray(m, o, d)
        eo, ed = np.zeros(3, np.float32), np.zeros(3, np.float32)
        R.ref_ray_transform(oracle._ptr(m), oracle._ptr(o), oracle._ptr(d), oracle._ptr(eo), oracle._ptr(ed))
        assert ro.tobytes() == eo.tobytes() and rd.tobytes() == ed.tobytes(), k


def test_affine_inverse(oracle):
    rng = np.random.default_rng(2)
    for _ in range(200):
        m = I.trs(rng.uniform(-5, 5, 3), rng.uniform(-3, 3, 3), rng.uniform(0.1, 4, 3))
        w = oracle.affine_inverse(m)
        a = np.vstack([m.reshape(3, 4), [0, 0, 0, 1]]).astype(np.float64)
        b = np.vstack([w.reshape(3, 4), [0, 0, 0, 1]]).astype(np.float64)
        assert np.allclose(b @ a, np.eye(4), atol=2e-6)
    # closed forms: a translation inverts to the negated translation, a
    # power-of-two scale to its reciprocal, both exactly
    w = oracle.affine_inverse(I.trs((1.5, -2.25, 3.0)))
    assert w.tolist() == [1, 0, 0, -1.5, 0, 1, 0, 2.25, 0, 0, 1, -3.0]
    w = oracle.affine_inverse(I.trs(s=(2.0, 0.5, 4.0)))
    assert w.tolist() == [0.5, 0, 0, 0, 0, 2.0, 0, 0, 0, 0, 0.25, 0]
    with pytest.raises(ValueError):
        oracle.affine_inverse(I.trs(s=(1.0, 0.0, 1.0)))


def test_world_box_bounds_transformed_vertices(oracle):
    sc = I.instanced_room()
    tl = oracle.TwoLevel(sc)
    inst = sc["instancing"]
    for i in range(len(inst["inst_mesh"])):
        m = inst["inst_mesh"][i]
        f, e = inst["mesh_first"][m], inst["mesh_first"][m + 1]
        pts = sc["tris"][f:e].reshape(-1, 3)
        w = np.array([oracle.xf_ray(inst["xform"][i], p, np.zeros(3, np.float32))[0] for p in pts])
        box = tl.boxes[i]
        assert (w >= box[:3]).all() and (w <= box[3:6]).all()
        # tight for axis-aligned placements (the identity room): faces touched
        if i == 0:
            assert np.array_equal(w.min(0), box[:3]) and np.array_equal(w.max(0), box[3:6])


def local_brute_force(oracle, sc, rays, tmin=0.0):
    """Closest hit over every instance's triangles with the ray in that
    instance's local space; earlier instances win exact ties."""
    inst = sc["instancing"]
    m = rays.shape[1]
    best_t = np.full(m, np.inf, np.float32)
    best = np.full(m, 0xFFFFFFFF, np.uint32)
    best_i = np.full(m, 0xFFFFFFFF, np.uint32)
    bary = np.zeros((2, m), np.float32)
    for i in range(len(inst["inst_mesh"])):
        w2l = oracle.affine_inverse(inst["xform"][i])
        lr = rays.copy()
        for k in range(m):
            o, d = oracle.xf_ray(w2l, rays[0:3, k], rays[3:6, k])
            lr[0:3, k], lr[3:6, k] = o, d
        lr[6] = np.minimum(rays[6], best_t)
        f, e = inst["mesh_first"][inst["inst_mesh"][i]], inst["mesh_first"][inst["inst_mesh"][i] + 1]
        p, t, b = oracle.trace_brute(sc["tris"][f:e], lr, tmin)
        hit = p != 0xFFFFFFFF
        best_t[hit], best[hit], best_i[hit] = t[hit], p[hit] + f, i
        bary[:, hit] = b[:, hit]
    return best_i, best, np.where(best == 0xFFFFFFFF, 0.0, best_t).astype(np.float32), bary


@pytest.mark.parametrize("width", [4, 2])
def test_two_level_trace_matches_local_brute_force(oracle, width):
    sc = I.instanced_room(n_spheres=7)
    tl = oracle.TwoLevel(sc)
    rays = I.rays_into(sc, 1500)
    gi, gp, gt, gb = tl.trace(rays, 0.0, width=width)
    ei, ep, et, eb = local_brute_force(oracle, sc, rays)
    assert (gp != 0xFFFFFFFF).mean() > 0.99
    assert np.array_equal(gt, et)
    assert np.array_equal(gp, ep) and np.array_equal(gi, ei)
    assert np.array_equal(gb, eb)
    # sphere instances are hit, under every kind of placement
    assert len(set(gi[gp != 0xFFFFFFFF].tolist())) == len(sc["instancing"]["inst_mesh"])


@pytest.mark.parametrize("width", [4, 2])
def test_identity_split_matches_single_level(oracle, width):
    base = N.gen_cornell(3)
    rays = I.rays_into(base, 2000, seed=8)
    ref_p, ref_t, ref_b, ref_st = oracle.LBVH(base["tris"]).trace(rays, 0.0, stats=True, width=width)
    for cuts in ([400, 900], [1]):
        tl = oracle.TwoLevel(I.split_identity(base, cuts))
        gi, gp, gt, gb = tl.trace(rays, 0.0, width=width)
        assert np.array_equal(gp, ref_p) and np.array_equal(gt, ref_t) and np.array_equal(gb, ref_b)
        assert (gi[gp != 0xFFFFFFFF] < len(cuts) + 1).all()
    # one mesh, one identity instance: the top tree is a single leaf, so the
    # counts are the mesh tree's own
    tl = oracle.TwoLevel(I.split_identity(base, []))
    _, _, _, _, st = tl.trace(rays, 0.0, stats=True, width=width)
    assert np.array_equal(st, ref_st)


@pytest.mark.parametrize("nee,mis", [(0, 0), (1, 1)])
def test_render_identity_split_matches_single_level(oracle, nee, mis):
    base = N.gen_cornell(3)
    two = I.split_identity(base, [300, 700, 1100])
    c = base["camera"]
    w, h = 40, 30
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render(base, oracle.LBVH(base["tris"]), cam, w, h, 2, 6, 3, 7, threads=8, nee=nee, mis=mis)
    got, seg2 = oracle.render(two, oracle.TwoLevel(two), cam, w, h, 2, 6, 3, 7, threads=8, nee=nee, mis=mis)
    assert seg == seg2 and np.array_equal(got, ref)


def test_instanced_light_list(oracle):
    sc = I.instanced_room()
    tl = oracle.TwoLevel(sc)
    nsph = int((sc["instancing"]["inst_mesh"] == 1).sum())
    # the room's two emissive triangles (instance 0), then both triangles of
    # each quad instance, made lights by their key override
    q0 = sc["instancing"]["mesh_first"][2]
    room_lights = np.nonzero((sc["prim_key"][:sc["instancing"]["mesh_first"][1]] >> 24) == 4)[0]
    assert tl.lights.tolist() == room_lights.tolist() + [q0, q0 + 1, q0, q0 + 1]
    assert tl.light_inst.tolist() == [0, 0, nsph + 1, nsph + 1, nsph + 2, nsph + 2]


def test_render_instanced_close_to_flattened(oracle):
    """Instancing changes rounding only: the flattened world-space scene
    renders to the same image up to noise."""
    sc = I.instanced_room()
    flat = I.flatten(sc)
    c = sc["camera"]
    w, h = 32, 24
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    a, _ = oracle.render(sc, oracle.TwoLevel(sc), cam, w, h, 6, 8, 3, 1, threads=8, nee=1, mis=1)
    b, _ = oracle.render(flat, oracle.LBVH(flat["tris"]), cam, w, h, 6, 8, 3, 1, threads=8, nee=1, mis=1)
    assert a.mean() > 0.05
    assert abs(a.mean() - b.mean()) / b.mean() < 0.02
    # most pixels agree closely (paths only part where a rounding flips a hit)
    close = np.isclose(a, b, rtol=1e-3, atol=1e-4).all(axis=2)
    assert close.mean() > 0.5
