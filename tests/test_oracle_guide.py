"""The oracle's path-guiding pieces pinned by mathematics and brute force.

- atan restatement against numpy's arctan;
- the concentric octahedral map: unit directions, round trip, and equal area
  (isotropic directions fall uniformly into the 64 bins);
- octree structure against a brute-force prefix construction, lookup against
  a brute-force deepest-prefix search;
- splat + filter against an independent numpy formulation;
- product sampler: pdf integrates to one, samples follow the pdf;
- guided rendering is unbiased (its mean matches a high-spp plain render).
"""
import numpy as np
import pytest

from meturay_amd import native as N


def iso_dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_atan_unit_matches_arctan(oracle):
    t = np.linspace(0.0, 1.0, 2001, dtype=np.float32)
    got = np.array([oracle.atan_unit(float(x)) for x in t])
    assert np.max(np.abs(got - np.arctan(t.astype(np.float64)))) < 3e-7


def test_octa_map_unit_and_round_trip(oracle):
    rng = np.random.default_rng(1)
    uv = rng.random((20000, 2)).astype(np.float32)
    d = oracle.octa_to_dirs(uv)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    back = oracle.dir_to_octas(d)
    # the map is continuous and one-to-one away from the seams of the
    # lower hemisphere fold; compare with the periodic distance there
    err = np.abs(back - uv)
    assert np.percentile(err, 99.9) < 1e-4
    # a known direction: +z is the centre of the map, -z its corners
    c = oracle.dir_to_octas(np.array([[0, 0, 1], [0, 0, -1]], dtype=np.float32))
    assert np.allclose(c[0], [0.5, 0.5])
    assert min(c[1][0], 1 - c[1][0]) < 1e-6 and min(c[1][1], 1 - c[1][1]) < 1e-6


def test_octa_bins_equal_area(oracle):
    n = 640000
    bins = oracle.dir_to_bins(iso_dirs(n, 2))
    h = np.bincount(bins, minlength=64)
    expect = n / 64
    chi2 = float(np.sum((h - expect) ** 2 / expect))
    # 63 degrees of freedom: chi2 above 120 has p < 1e-5
    assert chi2 < 120, chi2
    # bin centres map back to their own bins
    centres = np.array([[(b % 8 + 0.5) / 8, (b // 8 + 0.5) / 8] for b in range(64)], dtype=np.float32)
    assert np.array_equal(oracle.dir_to_bins(oracle.octa_to_dirs(centres)), np.arange(64))


def clustered_points(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.random((12, 3)) * 4 - 2
    pts = centres[rng.integers(0, 12, n)] + rng.normal(scale=0.15, size=(n, 3))
    return pts.astype(np.float32)


CUBE = np.array([-2.5, -2.5, -2.5, 5.0], dtype=np.float32)


def test_svo_codes_are_cells(oracle):
    pts = clustered_points(5000, 3)
    depth = 7
    codes = oracle.svo_codes(pts, CUBE, depth)
    cell = CUBE[3] / (1 << depth)
    q = np.clip(np.floor((pts - CUBE[:3]) / cell), 0, (1 << depth) - 1).astype(np.uint32)
    ref = np.array([oracle.lib().om_morton_compose30(int(a), int(b), int(c)) for a, b, c in q], dtype=np.uint32)
    # floor((p - lo) * inv_cell) in float32 vs the float64 division can differ
    # at exact cell boundaries only
    assert np.mean(codes == ref) > 0.999


@pytest.mark.parametrize("depth", [1, 4, 10])
def test_svo_structure_brute_force(oracle, depth):
    pts = clustered_points(3000, 4)
    codes = oracle.svo_codes(pts, CUBE, depth)
    s = oracle.Svo.build(codes, CUBE, depth)
    leaf = np.unique(codes)
    for l in range(depth + 1):
        lvl = np.unique(leaf >> np.uint32(3 * (depth - l)))
        o, c = s.level_offset[l], s.level_count[l]
        assert c == len(lvl)
        assert np.array_equal(s.codes[o:o + c], lvl)
    assert s.parent[0] == oracle.INVALID
    for g in range(1, s.nodes):
        p = s.parent[g]
        assert s.codes[p] == s.codes[g] >> 3
        assert (s.child_mask[p] >> (s.codes[g] & 7)) & 1
    for g in range(s.nodes):
        m = int(s.child_mask[g])
        kids = [k for k in range(s.nodes) if s.parent[k] == g] if s.nodes < 3000 else None
        if kids is not None:
            assert bin(m).count("1") == len(kids)
            if kids:
                assert s.first_child[g] == min(kids)
                assert kids == list(range(min(kids), min(kids) + len(kids)))
    # lookup: deepest existing prefix
    probe = np.concatenate([codes[:200], oracle.svo_codes(clustered_points(300, 9), CUBE, depth)])
    got = s.lookup(probe)
    node_of = {(l, int(c)): s.level_offset[l] + i for l in range(depth + 1)
               for i, c in enumerate(s.codes[s.level_offset[l]:s.level_offset[l] + s.level_count[l]])}
    for code, g in zip(probe, got):
        best = 0
        for l in range(depth + 1):
            k = (l, int(code) >> (3 * (depth - l)))
            if k in node_of:
                best = node_of[k]
            else:
                break
        assert g == best


def numpy_filter(s, codes, bins, values):
    """Independent restatement of splat + filter with numpy primitives."""
    n, depth = s.nodes, s.depth
    node = s.lookup(codes)
    fix = np.where(values > 0, np.floor(np.minimum(values, 32767.0).astype(np.float64) * 65536.0), 0)
    acc = np.zeros((n, 64), np.float64)
    cnt = np.zeros((n, 64), np.int64)
    np.add.at(acc, (node, bins), fix)
    np.add.at(cnt, (node, bins), 1)
    total = acc / 65536.0
    # up: every node holds its subtree's sums
    for l in range(depth, 0, -1):
        o, c = s.level_offset[l], s.level_count[l]
        np.add.at(total, s.parent[o:o + c], total[o:o + c])
        np.add.at(cnt, s.parent[o:o + c], cnt[o:o + c])
    mean = np.zeros_like(total)
    for l in range(depth + 1):
        o, c = s.level_offset[l], s.level_count[l]
        blk = slice(o, o + c)
        has = cnt[blk] > 0
        inherit = mean[s.parent[blk]] if l > 0 else np.zeros((c, 64))
        mean[blk] = np.where(has, total[blk] / np.maximum(cnt[blk], 1), inherit)
    grid = mean.reshape(n, 8, 8)
    pad = np.pad(grid, ((0, 0), (1, 1), (1, 1)))
    ones = np.pad(np.ones((8, 8)), 1)
    acc_b = sum(pad[:, 1 + dy:9 + dy, 1 + dx:9 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    k = sum(ones[1 + dy:9 + dy, 1 + dx:9 + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    rad = (acc_b / k).reshape(n, 64) + 1e-4
    return cnt, mean, rad


def test_svo_splat_filter_matches_numpy(oracle):
    depth = 6
    pts = clustered_points(20000, 5)
    codes = oracle.svo_codes(pts, CUBE, depth)
    s = oracle.Svo.build(codes[:15000], CUBE, depth)   # some splat codes miss the leaves
    rng = np.random.default_rng(6)
    bins = rng.integers(0, 64, len(codes)).astype(np.uint32)
    values = (rng.exponential(2.0, len(codes)) * (rng.random(len(codes)) < 0.7)).astype(np.float32)
    values[:5] = [0.0, -1.0, np.nan, 1e9, 32767.0]
    s.splat(codes, bins, values)
    s.filter()
    cnt, mean, rad = numpy_filter(s, codes, bins, np.nan_to_num(values, nan=0.0))
    assert np.array_equal(s.count, cnt.astype(np.uint32))
    assert np.array_equal(s.node_count, cnt.sum(axis=1).astype(np.uint32))
    assert np.allclose(s.sum, mean, rtol=2e-5, atol=1e-6)
    assert np.allclose(s.radiance, rad, rtol=2e-5, atol=1e-6)


def test_guide_pdf_normalised_and_sampled(oracle):
    depth = 3
    pts = clustered_points(4000, 7)
    codes = oracle.svo_codes(pts, CUBE, depth)
    s = oracle.Svo.build(codes, CUBE, depth)
    rng = np.random.default_rng(8)
    dirs = iso_dirs(len(codes), 9)
    bins = oracle.dir_to_bins(dirs)
    vals = (rng.exponential(1.0, len(codes)) * (1 + 20 * (bins == 27))).astype(np.float32)
    s.splat(codes, bins, vals)
    s.filter()
    node = int(s.lookup(codes[:1])[0])
    normal = np.array([0.3, -0.2, 0.93], dtype=np.float32)
    normal /= np.linalg.norm(normal)
    # piecewise-constant pdf over 64 equal-area bins: sum(pdf_b) * 4 pi / 64 = 1
    centres = oracle.octa_to_dirs(np.array([[(b % 8 + 0.5) / 8, (b // 8 + 0.5) / 8] for b in range(64)],
                                           dtype=np.float32))
    pb = s.pdf(np.full(64, node, np.uint32), np.tile(normal, (64, 1)), centres)
    assert abs(pb.sum() * 4 * np.pi / 64 - 1.0) < 1e-5
    w = s.radiance[node] * np.maximum(centres @ normal, 0)
    assert np.allclose(pb, w / w.sum() * 64 / (4 * np.pi), rtol=1e-5)
    # samples: bins drawn with probability pdf_b * area, pdf returned = pdf of the bin
    m = 200000
    xi = rng.random((m, 3)).astype(np.float32)
    d, pdf = s.sample(np.full(m, node, np.uint32), np.tile(normal, (m, 1)), xi)
    sb = oracle.dir_to_bins(d)
    h = np.bincount(sb, minlength=64)
    expect = pb * 4 * np.pi / 64 * m
    live = expect > 0
    assert h[~live].sum() <= m * 1e-4                     # only seam rounding
    chi2 = float(np.sum((h[live] - expect[live]) ** 2 / expect[live]))
    assert chi2 < live.sum() + 6 * np.sqrt(2 * live.sum()), chi2
    assert np.mean(np.isclose(pdf, pb[sb], rtol=1e-6)) > 0.999


def furnace(light_radiance):
    """Camera inside a closed emissive box looking at a Lambert icosphere
    (albedo 0.5).  With every wall at radiance 1 each sphere pixel's expected
    value is exactly 0.5; light_radiance > 1 makes the box's light patch
    brighter, so the learned guide is far from uniform."""
    sc = N.gen_cornell(3)
    keys = sc["prim_key"].copy()
    keys[:12] = (4 << 24) | 0
    keys[12:14] = (4 << 24) | 1
    keys[14:] = (1 << 24) | 0
    sc["prim_key"] = keys
    sc["emissive"] = np.array([[1, 1, 1], [light_radiance] * 3], np.float32)
    sc["lambert"] = np.array([[0.5, 0.5, 0.5]], np.float32)
    return sc


def test_guided_render_unbiased_furnace(oracle):
    sc = furnace(1.0)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 24, 24)
    g, seg, svo = oracle.render_guided(sc, bvh, cam, 24, 24, 8, 8, 1000, 9, 4, 5, 0.5, 64, threads=8,
                                       keep_svo=True)
    assert svo is not None and svo.nodes > 20 and seg > 0
    assert abs(g[10:14, 10:14].mean() - 0.5) < 0.01
    assert abs(g[0, 0].mean() - 1.0) < 1e-6
    svo.close()


def test_guided_render_unbiased_bright_light(oracle):
    sc = furnace(20.0)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 24, 24)
    ref, _ = oracle.render(sc, bvh, cam, 24, 24, 32, 8, 1000, 5, threads=8)
    g, _, _ = oracle.render_guided(sc, bvh, cam, 24, 24, 8, 8, 1000, 9, 4, 5, 0.5, 64, threads=8)
    centre = (slice(9, 15), slice(9, 15))
    assert abs(g[centre].mean() / ref[centre].mean() - 1.0) < 0.03


@pytest.mark.parametrize("mis", [0, 1])
def test_guided_render_with_nee_unbiased(oracle, mis):
    """Guiding with next-event estimation (and its MIS) estimates the same
    image as plain path tracing: the bright-light furnace, 3 %."""
    sc = furnace(20.0)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 24, 24)
    ref, _ = oracle.render(sc, bvh, cam, 24, 24, 32, 8, 1000, 5, threads=8)
    g, seg, svo = oracle.render_guided(sc, bvh, cam, 24, 24, 8, 8, 1000, 9, 4, 5, 0.5, 64, threads=8,
                                       keep_svo=True, nee=1, mis=mis)
    assert svo is not None and svo.nodes > 20
    centre = (slice(9, 15), slice(9, 15))
    assert abs(g[centre].mean() / ref[centre].mean() - 1.0) < 0.03
    svo.close()


def test_guided_nee_single_iteration_is_plain_nee(oracle):
    sc = N.gen_cornell(1)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 16, 16)
    a, sa = oracle.render(sc, bvh, cam, 16, 16, 2, 5, 2, 3, nee=1, mis=1)
    b, sb, _ = oracle.render_guided(sc, bvh, cam, 16, 16, 2, 5, 2, 3, 1, 5, nee=1, mis=1)
    assert sa == sb
    assert np.array_equal(a, b)


def test_guided_render_single_iteration_is_plain(oracle):
    sc = N.gen_cornell(1)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 16, 16)
    a, sa = oracle.render(sc, bvh, cam, 16, 16, 2, 5, 2, 3)
    b, sb, _ = oracle.render_guided(sc, bvh, cam, 16, 16, 2, 5, 2, 3, 1, 5)
    assert sa == sb
    assert np.array_equal(a, b)


def test_bin_centre_table_is_the_map(oracle):
    """The kernels' bin-centre table is exactly octa_to_dir of the centres."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "meturay_amd", "csrc", "mrh_guide.hpp")).read()
    body = src[src.index("GUIDE_CENTERS[GUIDE_BINS][3] = {"):]
    body = body[:body.index("};")]
    vals = [float.fromhex(t) for t in re.findall(r"([-+]?0x[0-9a-fA-F.]+p[-+]?\d+)f", body)]
    table = np.array(vals, dtype=np.float32).reshape(64, 3)
    uv = np.array([[(b % 8 + 0.5) / 8, (b // 8 + 0.5) / 8] for b in range(64)], dtype=np.float32)
    assert table.tobytes() == oracle.octa_to_dirs(uv).tobytes()
