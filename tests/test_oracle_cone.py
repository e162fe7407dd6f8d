"""Pins the oracle's cone trace through the guiding octree (om_svo_cone_trace;
AnisoSVOctreeGPU::ConeTraceRay, AnisoSVO.cuh:767-1032).

The reference's trace is device code with no host build, so it is pinned by
what the method fixes, against float64 brute force over the octree's boxes:
- a zero-aperture cone stopping only at leaves returns the first occupied
  leaf along the ray, entered at its slab entry t (or the ray's tmin when it
  starts inside), and a miss when no leaf lies on the ray;
- stopping `k` levels above the leaves returns the first existing node of
  that level along the ray;
- with an aperture, the node returned is one whose voxel the cone's disc is
  wider than at the t returned (or a leaf), and its box holds the ray point
  at that t.
Near ties (two boxes entered within 1e-5, grazing hits) are skipped."""
import numpy as np
import pytest

EPS = np.float32(1e-5)


def node_boxes(svo, cube, depth):
    """level, lo corner and size (world) of every node."""
    lo = np.array(cube[:3], np.float64)
    size = float(cube[3])
    lv = np.zeros(svo.nodes, np.int64)
    offs = list(svo.level_offset)
    for l in range(depth + 1):
        lv[offs[l]:offs[l + 1]] = l
    codes = svo.codes.astype(np.int64)

    def compact(v):
        v = v & 0x09249249
        v = (v | (v >> 2)) & 0x030C30C3
        v = (v | (v >> 4)) & 0x0300F00F
        v = (v | (v >> 8)) & 0x030000FF
        v = (v | (v >> 16)) & 0x000003FF
        return v

    xyz = np.stack([compact(codes), compact(codes >> 1), compact(codes >> 2)], 1).astype(np.float64)
    cell = size / (2.0 ** lv)
    return lv, lo + xyz * cell[:, None], cell


def slab(o, d, blo, bsz):
    inv = 1.0 / d
    t0 = (blo - o) * inv
    t1 = (blo + bsz[:, None] - o) * inv
    tn = np.minimum(t0, t1).max(1)
    tf = np.maximum(t0, t1).min(1)
    return tn, tf


def clamp_dir(d):
    d = d.astype(np.float32).copy()
    small = np.abs(d) < EPS
    d[small] = np.copysign(EPS, d[small])
    return d.astype(np.float64)


def make_tree(oracle, rng, depth, kind):
    cube = oracle.svo_cube([-1, -1, -1], [1, 1, 1], depth)
    if kind == "tris":
        c = rng.uniform(-0.8, 0.8, (40, 1, 3))
        tris = (c + 0.3 * rng.normal(size=(40, 3, 3))).clip(-1, 1).astype(np.float32).reshape(-1, 9)
        codes = oracle.voxelize(tris, cube, depth)
    else:
        pts = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
        codes = oracle.svo_codes(pts, cube, depth)
    return oracle.Svo.build(codes, cube.tolist(), depth), cube


def rays(rng, n):
    o = rng.uniform(-1.6, 1.6, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[::11] = np.eye(3)[rng.integers(0, 3, len(d[::11]))] * rng.choice([-1, 1], (len(d[::11]), 1))
    tmin = np.where(rng.random(n) < 0.3, rng.uniform(0, 0.5, n), 0.0)
    return np.hstack([o, d, tmin[:, None]]).astype(np.float32)


def first_box(o, d, tmin, blo, bsz):
    tn, tf = slab(o, d, blo, bsz)
    ok = (tf >= np.maximum(tn, tmin)) & (tf - np.maximum(tn, tmin) > 1e-6)
    if not ok.any():
        return None, None, None
    te = np.maximum(tn, tmin)
    te[~ok] = np.inf
    k = int(np.argmin(te))
    srt = np.sort(te)
    gap = srt[1] - srt[0] if np.isfinite(srt[1]) else np.inf
    return k, te[k], gap


@pytest.mark.parametrize("kind,depth", [("tris", 5), ("points", 4), ("tris", 6)])
def test_zero_aperture_hits_first_leaf(oracle, kind, depth):
    rng = np.random.default_rng(depth * 7 + len(kind))
    svo, cube = make_tree(oracle, rng, depth, kind)
    lv, blo, bsz = node_boxes(svo, cube, depth)
    leaves = np.where(lv == depth)[0]
    R = rays(rng, 400)
    nodes, t, leaf = svo.cone_trace(R, 0.0, 0)
    checked = hits = 0
    for i, r in enumerate(R):
        o, d, tmin = r[:3].astype(np.float64), clamp_dir(r[3:6]), float(r[6])
        k, te, gap = first_box(o, d, tmin, blo[leaves], bsz[leaves])
        if k is None:
            assert nodes[i] == 0xFFFFFFFF, i
            checked += 1
            continue
        if gap < 1e-5:
            continue
        checked += 1
        hits += 1
        assert leaf[i] == 1 and nodes[i] == leaves[k], (i, nodes[i], leaves[k])
        assert abs(float(t[i]) - te) < 1e-5 * max(1.0, te), (i, t[i], te)
    assert checked > 300 and hits > 40


@pytest.mark.parametrize("offset", [1, 2])
def test_offset_stops_at_level(oracle, offset):
    rng = np.random.default_rng(offset)
    depth = 6
    svo, cube = make_tree(oracle, rng, depth, "tris")
    lv, blo, bsz = node_boxes(svo, cube, depth)
    lvl = np.where(lv == depth - offset)[0]
    R = rays(rng, 800)
    nodes, t, leaf = svo.cone_trace(R, 0.0, offset)
    checked = 0
    for i, r in enumerate(R):
        o, d, tmin = r[:3].astype(np.float64), clamp_dir(r[3:6]), float(r[6])
        k, te, gap = first_box(o, d, tmin, blo[lvl], bsz[lvl])
        if k is None:
            assert nodes[i] == 0xFFFFFFFF
            continue
        if gap < 1e-5:
            continue
        checked += 1
        assert nodes[i] == lvl[k] and leaf[i] == 0
    assert checked > 100


@pytest.mark.parametrize("aperture", [1e-3, 2e-2, 0.3])
def test_aperture_stops_where_the_cone_covers_the_voxel(oracle, aperture):
    rng = np.random.default_rng(int(aperture * 1e4))
    depth = 7
    svo, cube = make_tree(oracle, rng, depth, "tris")
    lv, blo, bsz = node_boxes(svo, cube, depth)
    R = rays(rng, 300)
    nodes, t, leaf = svo.cone_trace(R, aperture, 0)
    leafsize = float(cube[3]) / (1 << depth)
    stops = 0
    for i, r in enumerate(R):
        if nodes[i] == 0xFFFFFFFF:
            continue
        n = int(nodes[i])
        L = int(lv[n])
        tt = float(t[i])
        diam2 = 4.0 * tt * tt * aperture / np.pi
        assert leaf[i] == (L == depth)
        if L < depth:
            stops += 1
            assert diam2 > (leafsize * 2 ** (depth - L)) ** 2 * (1 - 1e-5)
        p = r[:3].astype(np.float64) + tt * clamp_dir(r[3:6])
        tol = 1e-4 * float(cube[3])
        assert ((p >= blo[n] - tol) & (p <= blo[n] + bsz[n] + tol)).all(), (i, p, blo[n], bsz[n])
    if aperture >= 2e-2:
        assert stops > 20


def test_rays_leaving_the_cube_miss(oracle):
    rng = np.random.default_rng(3)
    svo, cube = make_tree(oracle, rng, 5, "points")
    R = np.array([[2, 2, 2, 1, 1, 1, 0], [-3, 0, 0, -1, 0, 0, 0], [0, 0, 0, 0, 0, 1, 5.0]], np.float32)
    nodes, t, leaf = svo.cone_trace(R, 0.0, 0)
    assert (nodes == 0xFFFFFFFF).all() and (leaf == 0).all()


def _room(oracle, depth):
    """A closed unit cube of 12 triangles, voxelised, and its octree."""
    c = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    quads = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (0, 3, 7, 4), (1, 2, 6, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris += [np.concatenate([c[a], c[b], c[cc]]), np.concatenate([c[a], c[cc], c[d]])]
    tris = np.array(tris, np.float32)
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)
    svo = oracle.Svo.build(oracle.voxelize(tris, cube, depth), cube.tolist(), depth)
    return svo, cube


def test_trace_fields_compose_the_cone_trace(oracle):
    depth, X, Y = 6, 16, 8
    svo, cube = _room(oracle, depth)
    rng = np.random.default_rng(9)
    n = svo.nodes
    irr = rng.uniform(0, 5, (n, 2)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    origins = np.hstack([rng.uniform(0.2, 0.8, (3, 3)), np.full((3, 1), 0.01)]).astype(np.float32)
    jit = rng.random((3, 2)).astype(np.float32)
    fields, nodes, t = svo.trace_fields(X, Y, origins, jit, irr, nrm, boundary=0.25)
    aperture = np.float32(4 * np.pi) / np.float32(X * X)
    for b in range(3):
        u = ((np.arange(X, dtype=np.float32)[None, :] + jit[b, 0]) / np.float32(X)).repeat(Y, 0)
        v = ((np.arange(Y, dtype=np.float32)[:, None] + jit[b, 1]) / np.float32(Y)).repeat(X, 1)
        z = oracle.octa_to_dirs(np.stack([u.ravel(), v.ravel()], 1))
        d = np.stack([z[:, 1], z[:, 2], z[:, 0]], 1)                  # y up
        rays = np.hstack([np.repeat(origins[b:b + 1, :3], X * Y, 0), d, np.repeat(origins[b:b + 1, 3:], X * Y, 0)])
        on, ot, _ = svo.cone_trace(rays, float(aperture), 0)
        assert np.array_equal(nodes[b].ravel(), on) and np.array_equal(t[b].ravel(), ot)
        for k in range(X * Y):
            want = np.float32(0.25) if on[k] == 0xFFFFFFFF else \
                irr[on[k], 0 if float(np.dot(nrm[on[k]], -d[k])) >= 0 else 1]
            assert fields[b].ravel()[k] == want


def test_closed_room_walls_fill_a_sixth_of_the_field_each(oracle):
    """From the centre of a cube room, each wall covers 4 pi / 6 of the sphere;
    the concentric octahedral map is equal-area, so about a sixth of the
    texels see each wall."""
    depth, X = 7, 64
    svo, cube = _room(oracle, depth)
    lv, blo, bsz = node_boxes(svo, cube, depth)
    ctr = blo + bsz[:, None] / 2
    dist = np.stack([ctr[:, 0], 1 - ctr[:, 0], ctr[:, 1], 1 - ctr[:, 1], ctr[:, 2], 1 - ctr[:, 2]], 1)
    wall = np.argmin(dist, 1).astype(np.float32) + 1
    irr = np.stack([wall, wall], 1)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (svo.nodes, 1))
    fields, nodes, _ = svo.trace_fields(X, X, np.array([[0.5, 0.5, 0.5, 0.0]], np.float32),
                                        np.array([[0.5, 0.5]], np.float32), irr, nrm)
    assert (nodes != 0xFFFFFFFF).all()
    frac = np.array([(fields == k).mean() for k in range(1, 7)])
    assert np.allclose(frac, 1 / 6, atol=0.02), frac


def _wall_tris(z=0.5, flip=False):
    a, b = 0.1, 0.9
    t = np.array([[a, a, z, b, a, z, b, b, z], [a, a, z, b, b, z, a, b, z]], np.float32)
    if flip:
        t = t.reshape(2, 3, 3)[:, ::-1].reshape(2, 9).copy()
    return t


def _cell_code(ix, iy, iz):
    def e(v):
        v &= 0x3FF
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v
    return e(ix) | (e(iy) << 1) | (e(iz) << 2)


def test_payload_flat_wall_sides_and_averages(oracle):
    """Wall z = 0.5 facing +z: every node's normal is +z (length 1); a vertex
    whose wo faces +z lands on the front side; leaf sides average their
    vertices, parents average their leaves, unsampled sides take the
    parent's, and a side unsampled everywhere keeps 65504."""
    depth = 5
    tris = _wall_tris()
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)
    codes, prims = oracle.voxelize(tris, cube, depth, prims=True)
    svo = oracle.Svo.build(codes, cube.tolist(), depth)
    L = svo.level_offset[depth]
    leaf = float(cube[3]) / (1 << depth)
    lv, blo, bsz = node_boxes(svo, cube, depth)
    rng = np.random.default_rng(2)
    leaves = np.arange(L, svo.nodes)
    pick = rng.choice(leaves, len(leaves) // 2, replace=False)        # half the leaves get samples
    verts, want = [], {}
    for lf in pick:
        ctr = blo[lf] + bsz[lf] / 2
        lum = rng.uniform(0.5, 4.0, 3).astype(np.float32)
        for x in lum:
            verts.append([*ctr, 0, 0, 1, x])
        want[int(lf)] = lum.astype(np.float64).mean()
    irr, nrm = svo.payload(tris, codes, prims, np.array(verts, np.float32))
    assert np.allclose(nrm[:, :3], [0, 0, 1]) and np.allclose(nrm[:, 3], 1.0)
    assert (irr[:, 1] == 65504.0).all()                              # nothing behind the wall
    for lf in leaves:
        if int(lf) in want:
            # 48.16 fixed point: each sample truncated to 2^-16
            assert abs(irr[lf, 0] - want[int(lf)]) < 2.0 ** -16 + 1e-6 * want[int(lf)]
    # parents of leaves: the mean of their sampled children; unsampled leaves
    # take their parent's value
    for nd in range(svo.level_offset[depth - 1], L):
        fc, k = svo.first_child[nd], bin(int(svo.child_mask[nd])).count("1")
        kids = [want[c] for c in range(fc, fc + k) if c in want]
        if kids:
            assert abs(irr[nd, 0] - np.mean(kids)) < 2.0 ** -16 + 1e-6 * np.mean(kids)
            for c in range(fc, fc + k):
                if c not in want:
                    assert irr[c, 0] == irr[nd, 0]


def test_payload_back_side_and_opposed_children(oracle):
    """Vertices whose wo points away from the normal go to the back side; a
    wall facing -z swaps its sides when averaged into a parent whose
    normal is +z."""
    depth = 4
    up = _wall_tris(0.25)
    down = _wall_tris(0.75, flip=True)                   # faces -z
    tris = np.concatenate([up, up, up, down])             # +z is the majority overall
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)
    codes, prims = oracle.voxelize(tris, cube, depth, prims=True)
    svo = oracle.Svo.build(codes, cube.tolist(), depth)
    lv, blo, bsz = node_boxes(svo, cube, depth)
    L = svo.level_offset[depth]
    verts = []
    for lf in range(L, svo.nodes):
        ctr = blo[lf] + bsz[lf] / 2
        verts.append([*ctr, 0, 0, 1, 2.0])                  # wo +z
    irr, nrm = svo.payload(tris, codes, prims, np.array(verts, np.float32))
    low = [lf for lf in range(L, svo.nodes) if blo[lf][2] < 0.5]
    high = [lf for lf in range(L, svo.nodes) if blo[lf][2] > 0.5]
    assert all(np.allclose(nrm[lf, :3], [0, 0, 1]) for lf in low)
    assert all(np.allclose(nrm[lf, :3], [0, 0, -1]) for lf in high)
    assert all(irr[lf, 0] == 2.0 for lf in low)            # wo . n >= 0: front
    assert all(irr[lf, 1] == 2.0 for lf in high)           # wo . n < 0: back
    # the root holds both walls; its normal follows the majority (+z) and the
    # -z wall's back side arrives on the root's front
    assert np.allclose(nrm[0, :3], [0, 0, 1]) and irr[0, 0] == 2.0


def test_payload_vertex_in_empty_cell_goes_to_nearest_leaf(oracle):
    depth = 4
    tris = _wall_tris(0.5)
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)
    codes, prims = oracle.voxelize(tris, cube, depth, prims=True)
    svo = oracle.Svo.build(codes, cube.tolist(), depth)
    leaf = float(cube[3]) / 16
    lv, blo, bsz = node_boxes(svo, cube, depth)
    lf = svo.level_offset[depth] + 5
    ctr = blo[lf] + bsz[lf] / 2
    p = ctr + np.array([0, 0, 0.8 * leaf])                  # in the empty cell above, near its floor
    irr, _ = svo.payload(tris, codes, prims, np.array([[*p, 0, 0, 1, 3.0]], np.float32))
    assert irr[lf, 0] == 3.0
    far = np.array([[0.5, 0.5, 0.95, 0, 0, 1, 7.0]], np.float32)   # no leaf among its neighbours
    irr2, _ = svo.payload(tris, codes, prims, far)
    assert (irr2 != 7.0).all()
