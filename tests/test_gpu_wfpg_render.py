"""The guided render with the field-traced guide of the voxel octree
(mrh_render_wfpg; WFPGTracer::Render / GenerateGuidedDirections / Finalize,
WFPGTracer.cu:590-760, 1590-1790).  Its stages are bit-exact against the
oracle one by one (test_gpu_vox.py, test_gpu_cone.py, test_gpu_wfpg.py,
test_gpu_guide.py for the ray binning); the render is checked for
unbiasedness against the oracle's 4096-spp converged images in
test_gpu_converged.py.  Here: one iteration is plain path tracing, the
voxel octree is the oracle's, the guided frames bin and sample rays, a
render is deterministic per path (equal path counts and bin counts when
repeated; the bin origin does not depend on the order of a bounce's rays),
and the argument checks."""
import numpy as np
import pytest

import instancing_scenes as I
from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_one_iteration_is_plain_path_tracing(ctx):
    scene = N.Scene(ctx, N.gen_cornell(2))
    a = torch.zeros(32 * 32 * 3, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    torch.cuda.synchronize()
    sa = scene.render(a.data_ptr(), 32, 32, 2, 6, 2, 4)
    sb, ws = scene.render_wfpg(b.data_ptr(), 32, 32, 2, 6, 2, 4, iterations=1)
    assert sa.segments == sb.segments
    assert torch.allclose(a, b, rtol=1e-5, atol=1e-6)
    assert ws.guided_bounces == 0 and ws.bins == 0
    scene.close()


@pytest.mark.parametrize("nee", [0, 1])
def test_guided_frames_bin_and_sample_deterministically(ctx, oracle, nee):
    sc = N.gen_cornell(2)
    w, h, S, depth, rr, seed, sd = 64, 48, 2, 6, 2, 5, 8
    print(f"[wfpg render] cornell {w}x{h} S={S} nee={nee}", flush=True)
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    kw = dict(iterations=3, svo_depth=sd, guide_ratio=0.5, min_bin_level=3, bin_ray_count=64, nee=nee, mis=nee)
    st, ws = scene.render_wfpg(img.data_ptr(), w, h, S, depth, rr, seed, **kw)
    first = img.cpu().numpy().copy()
    assert np.isfinite(first).all() and first.mean() > 0
    # the octree is the voxel octree of the scene's triangles in the
    # guide's cube, record for record and node for node
    tris = np.ascontiguousarray(sc["tris"], np.float32).reshape(-1, 9)
    p = tris.reshape(-1, 3)
    cube = oracle.svo_cube(p.min(0), p.max(0), sd)
    recs = oracle.voxelize(tris, cube, sd)
    assert ws.voxel_records == len(recs)
    o = oracle.Svo.build(recs, cube.tolist(), sd)
    assert ws.octree_nodes == o.nodes
    o.close()
    # two guided frames, each bounce with surface hits binned and sampled
    assert 2 <= ws.guided_bounces <= 2 * depth
    assert 0 < ws.bins <= ws.binned_rays
    assert ws.binned_rays < st.segments
    st2, ws2 = scene.render_wfpg(img.data_ptr(), w, h, S, depth, rr, seed, **kw)
    assert st2.segments == st.segments
    assert (ws2.bins, ws2.binned_rays, ws2.guided_bounces) == (ws.bins, ws.binned_rays, ws.guided_bounces)
    assert np.allclose(img.cpu().numpy(), first, rtol=1e-4, atol=1e-6)
    scene.close()


def test_product_mode_renders_the_same_image(ctx):
    """Product mode: the guide's half of the mixture samples field x BxDF.
    Its draws are hashed by the ray's place in the bounce, so only what does
    not depend on them is compared: the octree, and the image mean."""
    scene = N.Scene(ctx, N.gen_cornell(2))
    w, h = 64, 48
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    kw = dict(iterations=2, svo_depth=8, min_bin_level=3, bin_ray_count=64)
    st_m, ws_m = scene.render_wfpg(img.data_ptr(), w, h, 2, 6, 2, 9, **kw)
    mix = img.cpu().numpy().copy()
    st_p, ws_p = scene.render_wfpg(img.data_ptr(), w, h, 2, 6, 2, 9, product=1, **kw)
    prod = img.cpu().numpy()
    assert np.isfinite(prod).all() and prod.mean() > 0
    assert (ws_p.octree_nodes, ws_p.voxel_records) == (ws_m.octree_nodes, ws_m.voxel_records)
    assert ws_p.guided_bounces >= 2 and 0 < ws_p.bins <= ws_p.binned_rays < st_p.segments
    # both estimate the same image: the means agree to a few per cent at 4 spp x 2
    assert abs(prod.mean() / mix.mean() - 1.0) < 0.1
    with pytest.raises(N.MrhError):
        scene.render_wfpg(img.data_ptr(), w, h, 1, 4, 2, 1, product=2)
    scene.close()


def test_incremental_tracer_reproduces_the_render(ctx):
    """mrh_tracer_set_wfpg + `iterations` frames of the incremental tracer
    (the GPUTracerI protocol) = mrh_render_wfpg."""
    import ctypes as C
    sc = N.gen_cornell(2)
    scene = N.Scene(ctx, sc)
    w, h, S, depth, rr, seed = 48, 32, 2, 6, 2, 11
    wp = N.WfpgParams(3, 8, 0.5, 3, 64, 1.0, 0.0, 0)
    p = N.RenderParams(w, h, S, depth, rr, 0, seed, 0)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    ctx.check(N.lib().mrh_tracer_set_wfpg(t, C.byref(wp)))
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
    st_r, ws_r = scene.render_wfpg(ref.data_ptr(), w, h, S, depth, rr, seed, iterations=3, svo_depth=8,
                                   guide_ratio=0.5, min_bin_level=3, bin_ray_count=64)
    assert ws_r.guided_bounces > 0
    assert st.segments == st_r.segments
    assert torch.allclose(img, ref, rtol=1e-5, atol=1e-6)
    # a bad parameter set leaves no guide half set up
    bad = N.WfpgParams(3, 8, 0.5, 3, 0, 1.0, 0.0, 0)
    with pytest.raises(N.MrhError):
        ctx.check(N.lib().mrh_tracer_set_wfpg(t, C.byref(bad)))
    ctx.check(N.lib().mrh_tracer_destroy(t))
    scene.close()


def test_rejects_bad_arguments(ctx):
    scene = N.Scene(ctx, N.gen_cornell(2))
    b = torch.zeros(16 * 16 * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for kw in (dict(min_bin_level=11), dict(bin_ray_count=0), dict(bin_ray_count=4097), dict(filter_alpha=0.0),
               dict(filter_alpha=float("nan")), dict(boundary_radiance=-1.0), dict(iterations=0),
               dict(svo_depth=11), dict(guide_ratio=1.5)):
        with pytest.raises(N.MrhError):
            scene.render_wfpg(b.data_ptr(), 16, 16, 1, 4, 2, 1, **kw)
    scene.close()
    two = N.Scene(ctx, I.instanced_room(n_spheres=4))
    torch.cuda.synchronize()
    with pytest.raises(N.MrhError):
        two.render_wfpg(b.data_ptr(), 16, 16, 1, 4, 2, 1, iterations=2)
    two.close()
