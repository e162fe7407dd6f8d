"""GPU parity of the path-guiding octree and the guided tracer against the
oracle, through the C-ABI.

The guide is integer work (cells, octree links, counts, fixed-point sums)
plus float arithmetic in the oracle's operation order, so every array must
match bit for bit; the guided render must trace the same number of segments,
end with the identical guide, and give the oracle's image within relMSE 1e-4
(only the order of the per-pixel float atomics differs).
"""
import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CUBE = np.array([-2.5, -2.5, -2.5, 5.0], dtype=np.float32)


def dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def fetch(ctx, dptr, count, dtype):
    out = np.zeros(count, dtype)
    if count:
        ctx.memcpy(out, dptr, out.nbytes)
    return out


def clustered_points(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.random((12, 3)) * 4 - 2
    pts = centres[rng.integers(0, 12, n)] + rng.normal(scale=0.15, size=(n, 3))
    return pts.astype(np.float32)


def planes(a):
    """n x 3 -> 3 planes of n"""
    return np.ascontiguousarray(a.T)


def gpu_view(ctx, svo):
    v = svo.view()
    n = v.nodes
    return {
        "nodes": n,
        "level_count": np.array(v.level_count[:], np.uint32),
        "level_offset": np.array(v.level_offset[:], np.uint32),
        "codes": fetch(ctx, v.d_codes, n, np.uint32),
        "first_child": fetch(ctx, v.d_first_child, n, np.uint32),
        "child_mask": fetch(ctx, v.d_child_mask, n, np.uint32),
        "parent": fetch(ctx, v.d_parent, n, np.uint32),
        "count": fetch(ctx, v.d_count, n * 64, np.uint32).reshape(n, 64),
        "sum": fetch(ctx, v.d_sum, n * 64, np.float32).reshape(n, 64),
        "radiance": fetch(ctx, v.d_radiance, n * 64, np.float32).reshape(n, 64),
        "node_count": fetch(ctx, v.d_node_count, n, np.uint32),
    }


def assert_same_svo(g, o, filtered=True):
    assert g["nodes"] == o.nodes
    assert np.array_equal(g["level_count"], o.level_count)
    assert np.array_equal(g["level_offset"], o.level_offset)
    for k in ("codes", "first_child", "child_mask", "parent"):
        assert np.array_equal(g[k], getattr(o, k)), k
    if filtered:
        assert np.array_equal(g["count"], o.count)
        assert np.array_equal(g["node_count"], o.node_count)
        assert np.array_equal(g["sum"].view(np.uint32), o.sum.view(np.uint32))
        assert np.array_equal(g["radiance"].view(np.uint32), o.radiance.view(np.uint32))


@pytest.mark.parametrize("depth", [1, 5, 10])
def test_svo_codes_build_splat_filter_bit_exact(ctx, oracle, depth):
    n = 60000
    pts = clustered_points(n, 21)
    pts[:3] = [[-9, 0, 0], [9, 9, 9], [-2.5, -2.5, -2.5]]       # outside / on the cube border
    d_pts = dev(planes(pts))
    d_codes = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    svo = N.Svo(ctx)
    svo.codes(d_pts.data_ptr(), n, n, CUBE, depth, d_codes.data_ptr())
    ctx.sync()
    codes = host(d_codes, np.uint32)
    assert np.array_equal(codes, oracle.svo_codes(pts, CUBE, depth))

    nb = 45000                                                   # the rest splat into inner nodes
    svo.build(d_codes.data_ptr(), nb, CUBE, depth)
    o = oracle.Svo.build(codes[:nb], CUBE, depth)
    assert_same_svo(gpu_view(ctx, svo), o, filtered=False)

    rng = np.random.default_rng(22)
    bins = rng.integers(0, 64, n).astype(np.uint32)
    vals = (rng.exponential(3.0, n) * (rng.random(n) < 0.8)).astype(np.float32)
    vals[:4] = [np.nan, -2.0, 1e12, 32767.0]
    d_bins, d_vals = dev(bins), dev(vals)
    torch.cuda.synchronize()
    svo.splat(d_codes.data_ptr(), d_bins.data_ptr(), d_vals.data_ptr(), n)
    svo.filter()
    ctx.sync()
    o.splat(codes, bins, vals)
    o.filter()
    assert_same_svo(gpu_view(ctx, svo), o)

    # lookups, plain and with a sample threshold
    for mc in (0, 1, 40, 1000):
        d_nodes = torch.empty(n, dtype=torch.int32, device="cuda")
        svo.lookup(d_codes.data_ptr(), n, d_nodes.data_ptr(), mc)
        ctx.sync()
        got = host(d_nodes, np.uint32)
        ref = o.lookup(codes) if mc == 0 else o.lookup_guide(codes, mc)
        assert np.array_equal(got, ref), mc
    svo.close()
    o.close()


def test_svo_sample_pdf_bit_exact(ctx, oracle):
    depth = 4
    n = 50000
    pts = clustered_points(n, 31)
    codes = oracle.svo_codes(pts, CUBE, depth)
    rng = np.random.default_rng(32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    bins = oracle.dir_to_bins(dirs)
    vals = (rng.exponential(1.0, n) * (1 + 30 * (bins % 7 == 0))).astype(np.float32)
    o = oracle.Svo.build(codes, CUBE, depth)
    o.splat(codes, bins, vals)
    o.filter()
    svo = N.Svo(ctx)
    d_codes, d_bins, d_vals = dev(codes), dev(bins), dev(vals)
    torch.cuda.synchronize()
    svo.build(d_codes.data_ptr(), n, CUBE, depth)
    svo.splat(d_codes.data_ptr(), d_bins.data_ptr(), d_vals.data_ptr(), n)
    svo.filter()
    ctx.sync()
    m = 40000
    nodes = o.lookup(codes[:m])
    normals = rng.normal(size=(m, 3)).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    xi = rng.random((m, 3)).astype(np.float32)
    d_nodes, d_nrm, d_xi = dev(nodes), dev(planes(normals)), dev(planes(xi))
    d_dirs = torch.empty(3 * m, dtype=torch.float32, device="cuda")
    d_pdf = torch.empty(m, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    svo.sample(d_nodes.data_ptr(), d_nrm.data_ptr(), d_xi.data_ptr(), m, d_dirs.data_ptr(), d_pdf.data_ptr())
    ctx.sync()
    od, op = o.sample(nodes, normals, xi)
    gd = d_dirs.cpu().numpy().reshape(3, m).T
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(d_pdf.cpu().numpy().view(np.uint32), op.view(np.uint32))
    d_q = dev(planes(dirs[:m]))
    torch.cuda.synchronize()
    svo.pdf(d_nodes.data_ptr(), d_nrm.data_ptr(), d_q.data_ptr(), m, d_pdf.data_ptr())
    ctx.sync()
    assert np.array_equal(d_pdf.cpu().numpy().view(np.uint32), o.pdf(nodes, normals, dirs[:m]).view(np.uint32))
    svo.close()
    o.close()


def test_svo_errors(ctx):
    svo = N.Svo(ctx)
    d = torch.zeros(16, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(N.MrhError):
        svo.build(d.data_ptr(), 16, CUBE, 11)                    # depth out of range
    with pytest.raises(N.MrhError):
        svo.build(d.data_ptr(), 0, CUBE, 4)                      # no points
    with pytest.raises(N.MrhError):
        svo.filter()                                             # nothing built
    svo.close()


def _scene(name):
    if name == "cornell":
        return N.gen_cornell(2)
    return N.gen_sphere_grid(4, 8, 8, 5)


@pytest.mark.parametrize("name", ["cornell", "spheres"])
def test_guided_render_matches_oracle(ctx, oracle, name):
    sc = _scene(name)
    w, h, S, depth, rr, seed = 40, 30, 2, 6, 2, 3
    it, sd, ratio, ms = 3, 6, 0.5, 32
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    svo = N.Svo(ctx)
    torch.cuda.synchronize()
    st = scene.render_guided(img.data_ptr(), w, h, S, depth, rr, seed, iterations=it, svo_depth=sd,
                             guide_ratio=ratio, min_samples=ms, svo=svo)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg, osvo = oracle.render_guided(sc, bvh, cam, w, h, S, depth, rr, seed, it, sd, ratio, ms,
                                          threads=8, keep_svo=True)
    assert st.segments == seg
    assert_same_svo(gpu_view(ctx, svo), osvo)
    got = img.cpu().numpy().reshape(h, w, 3)
    err = float(np.mean((got - ref) ** 2 / (ref ** 2 + 1e-2)))
    assert err < 1e-4, err
    # deterministic: a second render reproduces the guide and the path count
    st2 = scene.render_guided(img.data_ptr(), w, h, S, depth, rr, seed, iterations=it, svo_depth=sd,
                              guide_ratio=ratio, min_samples=ms, svo=svo)
    assert st2.segments == st.segments
    assert_same_svo(gpu_view(ctx, svo), osvo)
    svo.close()
    osvo.close()
    scene.close()


def test_guided_render_one_iteration_equals_plain(ctx):
    sc = N.gen_cornell(2)
    scene = N.Scene(ctx, sc)
    a = torch.zeros(32 * 32 * 3, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    torch.cuda.synchronize()
    sa = scene.render(a.data_ptr(), 32, 32, 2, 6, 2, 4)
    sb = scene.render_guided(b.data_ptr(), 32, 32, 2, 6, 2, 4, iterations=1)
    assert sa.segments == sb.segments
    assert torch.allclose(a, b, rtol=1e-5, atol=1e-6)
    with pytest.raises(N.MrhError):
        scene.render_guided(b.data_ptr(), 32, 32, 2, 6, 2, 4, iterations=0)
    with pytest.raises(N.MrhError):
        scene.render_guided(b.data_ptr(), 32, 32, 2, 6, 2, 4, svo_depth=11)
    with pytest.raises(N.MrhError):
        scene.render_guided(b.data_ptr(), 32, 32, 2, 6, 2, 4, guide_ratio=1.5)
    scene.close()


def test_incremental_guided_tracer_matches_render_guided(ctx):
    import ctypes as C
    sc = N.gen_cornell(2)
    scene = N.Scene(ctx, sc)
    w, h, S, depth, rr, seed = 32, 24, 2, 6, 2, 7
    g = N.GuideParams(3, 6, 0.5, 16)
    p = N.RenderParams(w, h, S, depth, rr, 0, seed, 0)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    svo_t = N.Svo(ctx)
    ctx.check(N.lib().mrh_tracer_set_guide(t, C.byref(g), svo_t.h))
    cam = N.camera_desc(sc["camera"])
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
        more = C.c_int32(1)
        while more.value:
            ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    st = N.RenderStats()
    frames = C.c_uint32()
    ctx.check(N.lib().mrh_tracer_stats(t, C.byref(st), C.byref(frames)))
    assert frames.value == 3
    ref = torch.zeros_like(img)
    svo_r = N.Svo(ctx)
    st_r = scene.render_guided(ref.data_ptr(), w, h, S, depth, rr, seed, iterations=3, svo_depth=6,
                               guide_ratio=0.5, min_samples=16, svo=svo_r)
    assert st.segments == st_r.segments
    assert torch.allclose(img, ref, rtol=1e-5, atol=1e-6)
    a, b = gpu_view(ctx, svo_t), gpu_view(ctx, svo_r)
    for k in ("codes", "count", "radiance"):
        assert np.array_equal(a[k], b[k]), k
    ctx.check(N.lib().mrh_tracer_destroy(t))
    svo_t.close()
    svo_r.close()
    scene.close()


@pytest.mark.parametrize("mis", [0, 1])
def test_guided_render_with_nee_matches_oracle(ctx, oracle, mis):
    """Guiding with next-event estimation: the oracle's segments (path and
    next-event rays), its octree bit for bit, and its image."""
    import nee_scenes as S
    sc = S.lit_spheres()
    w, h, S_, depth, rr, seed = 40, 30, 2, 8, 3, 6
    it, sd, ratio, ms = 3, 6, 0.5, 32
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    svo = N.Svo(ctx)
    torch.cuda.synchronize()
    st = scene.render_guided(img.data_ptr(), w, h, S_, depth, rr, seed, iterations=it, svo_depth=sd,
                             guide_ratio=ratio, min_samples=ms, svo=svo, nee=1, mis=mis)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg, osvo = oracle.render_guided(sc, bvh, cam, w, h, S_, depth, rr, seed, it, sd, ratio, ms, threads=8,
                                          keep_svo=True, tree_width=ctx.tune_get("trace_width"), nee=1, mis=mis)
    assert st.segments == seg
    assert_same_svo(gpu_view(ctx, svo), osvo)
    got = img.cpu().numpy().reshape(h, w, 3)
    assert ref.mean() > 0.01
    err = float(np.mean((got - ref) ** 2 / (ref ** 2 + 1e-2)))
    assert err < 1e-4, err
    svo.close()
    osvo.close()
    scene.close()
