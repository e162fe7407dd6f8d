"""Pins the oracle's triangle voxeliser (oracle/om_vox.c; the reference's
TriFunctions::Voxelize, GPUPrimitiveTriangle.h:384-620).

The reference's voxeliser is device code with no host build, so it leaves no
golden vectors; the oracle is pinned by what the method fixes, checked in
float64 per triangle:
- conservative coverage: every pixel of the raster plane whose square
  overlaps the projected triangle (robustly, by a 1e-3 pixel margin) gets a
  record, and no record lies outside the triangle's pixel rectangle;
- one record per covered pixel (no pixel twice within a triangle);
- the raster plane is the one of the normal's dominant axis, and each
  record's index along that axis is the leaf of the triangle's plane point
  under the pixel centre (within one leaf: the float plane point can round
  across a leaf face);
- the cube of the reference (scene bounds padded by one leaf, AnisoSVO.cu:
  670-678) and known answers for an axis-aligned square."""
import numpy as np
import pytest


def compact10(v):
    v = np.asarray(v, np.uint64) & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    v = (v | (v >> 16)) & 0x000003FF
    return v.astype(np.int64)


def decode(codes):
    c = np.asarray(codes, np.uint64)
    return np.stack([compact10(c), compact10(c >> np.uint64(1)), compact10(c >> np.uint64(2))], 1)


def raster_axes(tri):
    p = tri.reshape(3, 3).astype(np.float64)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    a = np.abs(n)
    dom = int(np.argmax(a))
    srt = np.sort(a)
    return dom, [k for k in range(3) if k != dom], (srt[2] - srt[1]) > 1e-6 * srt[2]


def tri_overlaps_square(t2, x0, y0, x1, y1):
    """2D separating-axis test of a triangle and an axis-aligned box (float64)."""
    box = np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]])
    if t2[:, 0].max() < x0 or t2[:, 0].min() > x1 or t2[:, 1].max() < y0 or t2[:, 1].min() > y1:
        return False
    for k in range(3):
        e = t2[(k + 1) % 3] - t2[k]
        nrm = np.array([-e[1], e[0]])
        pt = t2 @ nrm
        pb = box @ nrm
        if pt.max() < pb.min() or pt.min() > pb.max():
            return False
    return True


def check_triangle(oracle, tri, cube, depth):
    R = 1 << depth
    codes = oracle.voxelize(tri, cube, depth)
    vox = decode(codes)
    dom, (ua, va), clear = raster_axes(tri)
    if not clear:
        return None                                  # near tie of the dominant axis: skip
    p = tri.reshape(3, 3).astype(np.float64)
    lo, size = np.array(cube[:3], np.float64), float(cube[3])
    g = (p - lo) / size * R                          # leaf-grid coordinates
    pix = set(zip(vox[:, ua].tolist(), vox[:, va].tolist()))
    assert len(pix) == len(codes), "a pixel gave two records"
    t2 = g[:, [ua, va]]
    xr = (int(np.floor(t2[:, 0].min() - 1e-4)), int(np.ceil(t2[:, 0].max() + 1e-4)))
    yr = (int(np.floor(t2[:, 1].min() - 1e-4)), int(np.ceil(t2[:, 1].max() + 1e-4)))
    for (x, y) in pix:
        assert xr[0] <= x < xr[1] and yr[0] <= y < yr[1], (x, y, xr, yr)
    m = 1e-3
    for x in range(max(0, xr[0]), min(R, xr[1])):
        for y in range(max(0, yr[0]), min(R, yr[1])):
            if tri_overlaps_square(t2, x + m, y + m, x + 1 - m, y + 1 - m):
                assert (x, y) in pix, ("overlapping pixel not covered", x, y)
    # the dominant-axis index: the plane point under the pixel centre
    nrm = np.cross(g[1] - g[0], g[2] - g[0])
    for (vx, x, y) in zip(vox, vox[:, ua], vox[:, va]):
        cu, cv = x + 0.5, y + 0.5
        depth_at = (nrm @ g[0] - nrm[ua] * cu - nrm[va] * cv) / nrm[dom]
        want = min(max(int(np.floor(depth_at)), 0), R - 1)
        assert abs(int(vx[dom]) - want) <= 1, (vx, depth_at)
    return len(codes)


def test_cube_pads_one_leaf(oracle):
    c = oracle.svo_cube([0, 0, 0], [2, 1, 1], 4)
    assert np.allclose(c, [-0.125, -0.125, -0.125, 2.25])


def test_axis_aligned_square_known_answer(oracle):
    depth, R = 4, 16
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)     # leaf = 1.125 / 16
    z = 0.5
    a, b = 0.25, 0.75
    tris = np.array([[a, a, z, b, a, z, b, b, z], [a, a, z, b, b, z, a, b, z]], np.float32)
    vox = decode(oracle.voxelize(tris, cube, depth))
    assert (vox[:, 2] == 8).all()                          # (0.5 + 1/16) / 1.125 * 16 = 8 exactly
    inside = {(x, y) for x in range(4, 12) for y in range(4, 12)}   # pixels the square overlaps
    got = set(zip(vox[:, 0].tolist(), vox[:, 1].tolist()))
    assert inside <= got
    # conservative extras only next to the square's corners (the 45-degree
    # corners of the two halves reach past the square)
    assert all(3 <= x <= 12 and 3 <= y <= 12 for x, y in got - inside)


@pytest.mark.parametrize("depth", [3, 5, 7])
def test_random_triangles_properties(oracle, depth):
    rng = np.random.default_rng(depth)
    cube = oracle.svo_cube([-1, -1, -1], [1, 1, 1], depth)
    checked = 0
    for i in range(120):
        c = rng.uniform(-0.9, 0.9, 3)
        s = rng.choice([0.02, 0.2, 0.8])
        tri = (c + s * rng.normal(size=(3, 3))).clip(-1, 1).astype(np.float32).reshape(-1)
        if i % 10 == 0:
            tri.reshape(3, 3)[:, i % 3] = tri[i % 3]       # axis-aligned plane
        if check_triangle(oracle, tri, cube, depth) is not None:
            checked += 1
    assert checked > 100


def test_degenerate_triangles_emit_nothing(oracle):
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], 6)
    line = np.array([0.1, 0.1, 0.1, 0.5, 0.5, 0.5, 0.9, 0.9, 0.9], np.float32)
    point = np.array([0.3] * 9, np.float32)
    assert len(oracle.voxelize(line, cube, 6)) == 0
    assert len(oracle.voxelize(point, cube, 6)) == 0


def test_records_follow_triangle_order(oracle):
    rng = np.random.default_rng(4)
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], 6)
    tris = rng.uniform(0.1, 0.9, (20, 9)).astype(np.float32)
    allc = oracle.voxelize(tris, cube, 6)
    parts = [oracle.voxelize(t, cube, 6) for t in tris]
    assert np.array_equal(allc, np.concatenate(parts))
