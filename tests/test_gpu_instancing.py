"""Two-level scenes on the GPU, through the C-ABI, against the oracle.

Ray casts (mrh_scene_trace_closest) return the oracle's instance and
primitive ids, bit-identical t and barycentrics, and the same node / test
counts (top tree plus mesh trees) with both mesh-tree widths.  Renders
trace the same segments and match the oracle image within relMSE 1e-4 (the
per-pixel sums are float atomics), with and without next-event estimation.
"""
import numpy as np
import pytest

import instancing_scenes as I
from meturay_amd import native as N

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


@pytest.fixture(params=[4, 2])
def width(request, ctx):
    old = ctx.tune_get("trace_width")
    ctx.tune("trace_width", request.param)
    yield request.param
    ctx.tune("trace_width", old)


def gpu_trace(ctx, scene, rays, tmin=0.0):
    m = rays.shape[1]
    dr = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    inst = torch.empty(m, dtype=torch.int32, device="cuda")
    prim = torch.empty(m, dtype=torch.int32, device="cuda")
    t = torch.empty(m, dtype=torch.float32, device="cuda")
    bary = torch.empty(2 * m, dtype=torch.float32, device="cuda")
    st = torch.empty(2 * m, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    scene.trace_closest(dr, m, m, tmin, inst, prim, t, bary, st)
    u = lambda x: x.cpu().numpy().view(np.uint32)
    return (u(inst), u(prim), t.cpu().numpy(), bary.cpu().numpy().reshape(2, m),
            u(st).reshape(m, 2))


def test_scene_trace_two_level_matches_oracle(ctx, oracle, width):
    sc = I.instanced_room(n_spheres=9)
    scene = N.Scene(ctx, sc)
    rays = I.rays_into(sc, 20000)
    gi, gp, gt, gb, gs = gpu_trace(ctx, scene, rays)
    ei, ep, et, eb, es = oracle.TwoLevel(sc).trace(rays, 0.0, stats=True, width=width)
    assert (ep != 0xFFFFFFFF).mean() > 0.99
    assert np.array_equal(gp, ep) and np.array_equal(gi, ei)
    assert np.array_equal(gt, et) and np.array_equal(gb, eb)
    assert np.array_equal(gs, es)
    scene.close()


def test_scene_trace_single_level_matches_oracle(ctx, oracle, width):
    sc = N.gen_cornell(3)
    scene = N.Scene(ctx, sc)
    rays = I.rays_into(sc, 5000, seed=4)
    gi, gp, gt, gb, gs = gpu_trace(ctx, scene, rays)
    ep, et, eb, es = oracle.LBVH(sc["tris"]).trace(rays, 0.0, stats=True, width=width)
    assert np.array_equal(gp, ep) and np.array_equal(gt, et) and np.array_equal(gb, eb)
    assert np.array_equal(gs, es)
    assert np.array_equal(gi, np.where(ep == 0xFFFFFFFF, 0xFFFFFFFF, 0).astype(np.uint32))
    scene.close()


def _render_pair(ctx, oracle, sc, w, h, S, depth, rr, seed, nee=0, mis=0):
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), w, h, S, depth, rr, seed, nee=nee, mis=mis)
    g = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render(sc, oracle.TwoLevel(sc), cam, w, h, S, depth, rr, seed, threads=8,
                             tree_width=ctx.tune_get("trace_width"), nee=nee, mis=mis)
    return g, st, ref, seg


@pytest.mark.parametrize("nee,mis", [(0, 0), (1, 0), (1, 1)])
def test_render_two_level_matches_oracle(ctx, oracle, width, nee, mis):
    sc = I.instanced_room()
    g, st, ref, seg = _render_pair(ctx, oracle, sc, 64, 48, 2, 10, 3, 5, nee, mis)
    assert st.segments == seg
    assert ref.mean() > 0.02
    assert relmse(g, ref) <= 1e-4


def test_render_identity_split_equals_single_level(ctx):
    base = N.gen_cornell(3)
    two = I.split_identity(base, [300, 700, 1100])
    imgs = []
    for sc in (base, two):
        scene = N.Scene(ctx, sc)
        img = torch.zeros(48 * 64 * 3, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        st = scene.render(img.data_ptr(), 64, 48, 2, 8, 3, 9, nee=1, mis=1)
        imgs.append((img.cpu().numpy(), st.segments))
        scene.close()
    assert imgs[0][1] == imgs[1][1]
    assert np.allclose(imgs[0][0], imgs[1][0], rtol=1e-5, atol=1e-6)


def test_two_level_errors(ctx):
    sc = I.instanced_room(n_spheres=2)

    def bad(**kw):
        s = dict(sc)
        inst = dict(sc["instancing"])
        inst.update(kw)
        s["instancing"] = inst
        with pytest.raises(N.MrhError):
            N.Scene(ctx, s)

    xf = sc["instancing"]["xform"].copy()
    xf[1, :] = 0.0
    bad(xform=xf)                                                    # singular
    xf = sc["instancing"]["xform"].copy()
    xf[2, 3] = np.inf
    bad(xform=xf)                                                    # not finite
    first = sc["instancing"]["mesh_first"].copy()
    first[-1] += 1
    bad(mesh_first=first)                                            # ranges miss the triangles
    first = sc["instancing"]["mesh_first"].copy()
    first[2] = first[1]
    bad(mesh_first=first)                                            # empty mesh
    im = sc["instancing"]["inst_mesh"].copy()
    im[0] = 3
    bad(inst_mesh=im)                                                # no such mesh
    ik = sc["instancing"]["inst_key"].copy()
    ik[1] = N.key(4, 7)
    bad(inst_key=ik)                                                 # no such material
    scene = N.Scene(ctx, sc)
    with pytest.raises(N.MrhError):
        scene.bvh()                                                  # no single LBVH
    scene.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_engine_mixes_two_level_and_single_level(ctx, graphs):
    """Two-level and single-level requests traced together in the serving
    engine give their solo segments and images."""
    two = N.Scene(ctx, I.instanced_room())
    one = N.Scene(ctx, N.gen_cornell(2))
    # (scene, w, h, strata, depth, rr, seed, iterations, nee, mis)
    mix = [(two, 48, 32, 2, 8, 3, 1, 1, 0, 0), (one, 40, 30, 2, 6, 2, 2, 1, 0, 0),
           (two, 32, 32, 1, 10, 3, 3, 1, 1, 1), (two, 40, 24, 2, 8, 3, 4, 2, 0, 0),
           (one, 24, 24, 1, 6, 3, 5, 1, 1, 0), (two, 24, 16, 2, 6, 2, 6, 1, 1, 0)]
    ref = []
    for sc, w, h, S, dep, rr, seed, it, nee, mis in mix:
        img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        if it > 1:
            st = sc.render_guided(img.data_ptr(), w, h, S, dep, rr, seed, iterations=it, svo_depth=6,
                                  guide_ratio=0.5, min_samples=16)
        else:
            st = sc.render(img.data_ptr(), w, h, S, dep, rr, seed, nee=nee, mis=mis)
        ref.append((img.cpu().numpy(), st.segments))
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=4, use_graphs=graphs)
    imgs = [torch.full((m[1] * m[2] * 3,), 7.0, dtype=torch.float32, device="cuda") for m in mix]
    torch.cuda.synchronize()
    for k, (sc, w, h, S, dep, rr, seed, it, nee, mis) in enumerate(mix):
        server.submit(sc, imgs[k], w, h, S, dep, rr, seed, iterations=it, svo_depth=6, guide_ratio=0.5,
                      min_samples=16, tag=k, nee=nee, mis=mis)
    done = server.run()
    assert sorted(d["tag"] for d in done) == list(range(len(mix)))
    for d in done:
        k = d["tag"]
        assert d["status"] == 0 and d["segments"] == ref[k][1], (k, d["segments"], ref[k][1])
        assert np.allclose(imgs[k].cpu().numpy(), ref[k][0], rtol=1e-4, atol=1e-6), k
    assert server.stats().steps > 10
    server.close()
    two.close()
    one.close()


def test_scene_file_instanced_render_matches_oracle(ctx, oracle):
    """A scene file loaded as a two-level scene (MRH_SCENE_FILE_INSTANCED)
    renders as the oracle's two-level restatement does."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "scenes", "room.json")
    sc = N.load_scene_file(path, instanced=True)
    g, st, ref, seg = _render_pair(ctx, oracle, sc, 48, 48, 2, 8, 3, 2, 1, 1)
    assert st.segments == seg
    assert ref.mean() > 0.01
    assert relmse(g, ref) <= 1e-4


def test_guided_render_two_level_matches_oracle(ctx, oracle):
    """Path guiding over a two-level scene: the octree is built in the top
    tree's world bounds from world-space path vertices, as in the oracle."""
    sc = I.instanced_room()
    w, h, S, depth, rr, seed = 40, 30, 2, 6, 2, 3
    it, sd, ratio, ms = 3, 6, 0.5, 32
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render_guided(img.data_ptr(), w, h, S, depth, rr, seed, iterations=it, svo_depth=sd,
                             guide_ratio=ratio, min_samples=ms)
    got = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg, _ = oracle.render_guided(sc, oracle.TwoLevel(sc), cam, w, h, S, depth, rr, seed, it, sd, ratio, ms,
                                       threads=8, tree_width=ctx.tune_get("trace_width"))
    assert st.segments == seg
    assert ref.mean() > 0.02
    assert relmse(got, ref) <= 1e-4


def test_filtered_render_two_level_matches_oracle(ctx, oracle):
    """A reconstruction-filtered frame of a two-level scene keeps every
    sample's radiance exactly, so it equals the oracle's filtered frame."""
    sc = I.instanced_room()
    w, h, S = 48, 36, 2
    d = N.filter_desc("Tent", 1.0)
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render_filtered(img.data_ptr(), w, h, S, 8, 3, 4, filt=d)
    got = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render_filtered(sc, oracle.TwoLevel(sc), cam, w, h, S, 8, 3, 4,
                                      oracle.OmFilterDesc(d.type, d.radius, d.alpha, d.b, d.c), threads=8,
                                      tree_width=ctx.tune_get("trace_width"))
    assert st.segments == seg
    assert np.array_equal(got, ref)


def test_incremental_tracer_two_level_matches_render(ctx):
    """mrh_tracer_* (GenerateWork / Render per bounce / Finalize) over a
    two-level scene traces the same segments as mrh_render."""
    import ctypes as C
    sc = I.instanced_room()
    scene = N.Scene(ctx, sc)
    w, h = 40, 30
    ref = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(ref.data_ptr(), w, h, 2, 8, 3, 0)
    p = N.RenderParams(w, h, 2, 8, 3, 0, 0, 0, 0)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    cam = N.camera_desc(sc["camera"])
    img = torch.zeros_like(ref)
    ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
    more = C.c_int32(1)
    while more.value:
        ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    ts = N.RenderStats()
    frames = C.c_uint32()
    ctx.check(N.lib().mrh_tracer_stats(t, C.byref(ts), C.byref(frames)))
    ctx.check(N.lib().mrh_tracer_destroy(t))
    assert ts.segments == st.segments
    assert np.allclose(img.cpu().numpy(), ref.cpu().numpy(), rtol=1e-5, atol=1e-6)
    scene.close()


@pytest.mark.parametrize("bits", [30, 63])
def test_two_level_edge_shapes(ctx, oracle, bits):
    """Edge shapes: one-triangle meshes (a leaf root), one instance (a leaf
    top tree), a mesh used by many instances and 63-bit codes."""
    rng = np.random.default_rng(21)
    one = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    blob = I.uv_sphere(6, 8, 1.0)
    tris = np.concatenate([one, blob]).astype(np.float32)
    keys = np.full(len(tris), N.key(1, 0), np.uint32)
    base = {"tris": tris, "prim_key": keys, "lambert": np.array([[0.5, 0.5, 0.5]], np.float32),
            "metal": np.zeros((0, 10), np.float32), "dielectric": np.zeros((0, 4), np.float32),
            "emissive": np.zeros((0, 3), np.float32), "reflect": np.zeros((0, 4), np.float32)}
    first = np.array([0, 1, len(tris)], np.uint32)
    cases = {
        "single instance of the triangle": ([I.trs((0.1, 0.2, 0.3), (0.3, 0.1, 0.2), (2, 2, 2))], [0]),
        "single instance of the blob": ([I.trs((0, 0, 0), (0.5, 0.2, 0.1), (0.5, 0.7, 0.9))], [1]),
        "many instances": ([I.trs(rng.uniform(-0.8, 0.8, 3), rng.uniform(-3, 3, 3), rng.uniform(0.05, 0.2, 3))
                            for _ in range(40)], [int(k % 2) for k in range(40)]),
    }
    for name, (xf, mesh) in cases.items():
        sc = dict(base)
        sc["instancing"] = {"mesh_first": first, "xform": np.stack(xf).astype(np.float32),
                            "inst_mesh": np.array(mesh, np.uint32)}
        scene = N.Scene(ctx, sc, morton_bits=bits)
        rays = I.rays_into(sc, 4000, seed=7)
        gi, gp, gt, gb, gs = gpu_trace(ctx, scene, rays)
        ei, ep, et, eb, es = oracle.TwoLevel(sc, bits).trace(rays, 0.0, stats=True,
                                                              width=ctx.tune_get("trace_width"))
        assert (ep != 0xFFFFFFFF).any(), name
        assert np.array_equal(gp, ep) and np.array_equal(gi, ei), name
        assert np.array_equal(gt, et) and np.array_equal(gb, eb), name
        assert np.array_equal(gs, es), name
        scene.close()
