"""Next-event estimation and direct-light MIS on the GPU, through the C-ABI,
against the oracle rendered from the same seed.

Every path draws the same numbers on both sides, so the segment counts
(path and next-event rays) must be equal and the images within relMSE 1e-4
(they differ only in the order of the per-pixel float additions).  Scenes:
the Cornell box, the sphere grid with two spheres turned into lights (all
three BxDFs), the emissive furnace, a light hidden behind another light,
and a mirror under a light (specular vertices send no NEE ray).
"""
import ctypes as C

import numpy as np
import pytest

from meturay_amd import native as N
import nee_scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


SCENES = {
    "cornell": (lambda: N.gen_cornell(3), 64, 48, 2, 8, 3),
    "lit_spheres": (S.lit_spheres, 96, 64, 2, 10, 3),
    "furnace": (lambda: S.furnace(1.0), 32, 32, 2, 8, 1000),
    "hidden_light": (lambda: S.square_lights_scene([(1.0, 0.5, 2.0), (2.0, 0.5, 5.0)]), 32, 32, 2, 2, 10),
    "mirror": (lambda: S.square_lights_scene([(1.0, 0.5, 2.0)], mirror=True), 16, 16, 2, 4, 10),
    "reflect_spheres": (lambda: S.lit_spheres(reflect=True), 96, 64, 2, 10, 3),
}


def _pair(ctx, oracle, sc, w, h, spp_axis, depth, rr, seed, nee, mis):
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), w, h, spp_axis, depth, rr, seed, nee=nee, mis=mis)
    g = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    ob = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render(sc, ob, cam, w, h, spp_axis, depth, rr, seed, threads=8,
                             tree_width=ctx.tune_get("trace_width"), nee=nee, mis=mis)
    return g, st, ref, seg


@pytest.mark.parametrize("mis", [0, 1])
@pytest.mark.parametrize("name", list(SCENES))
def test_render_nee_matches_oracle(ctx, oracle, name, mis):
    make, w, h, spp_axis, depth, rr = SCENES[name]
    g, st, ref, seg = _pair(ctx, oracle, make(), w, h, spp_axis, depth, rr, 11, 1, mis)
    assert st.segments == seg
    assert ref.mean() > 0.01
    assert relmse(g, ref) <= 1e-4
    assert np.abs(g - ref).max() <= 1e-4 * max(1.0, ref.max())


def test_render_reflect_without_nee(ctx, oracle):
    g, st, ref, seg = _pair(ctx, oracle, S.lit_spheres(reflect=True), 96, 64, 2, 10, 3, 5, 0, 0)
    assert st.segments == seg
    assert relmse(g, ref) <= 1e-4
    a, le = (0.9, 0.7, 0.5), 2.0
    g, st, ref, seg = _pair(ctx, oracle, S.square_lights_scene([(1.0, 0.5, le)], reflect=(a, 0.0)), 16, 16, 2, 4,
                            10, 3, 0, 0)
    want = np.array([np.float32(le) * np.float32(x) for x in a], np.float32)
    # four equal samples per pixel, summed in float then scaled: within an ulp
    assert np.allclose(g, np.broadcast_to(want, g.shape), rtol=1e-6, atol=0)
    assert st.segments == seg == 16 * 16 * 4 * 2


def test_nee_adds_next_event_rays(ctx, oracle):
    sc = N.gen_cornell(3)
    _, plain, _, _ = _pair(ctx, oracle, sc, 48, 48, 1, 6, 3, 2, 0, 0)
    _, nee, _, seg = _pair(ctx, oracle, sc, 48, 48, 1, 6, 3, 2, 1, 0)
    assert nee.segments == seg and nee.segments > plain.segments
    # next-event rays ride along with the next bounce's path rays
    assert nee.bounce_rays[0] == plain.bounce_rays[0] == 48 * 48
    assert nee.bounce_rays[1] > plain.bounce_rays[1]


def test_incremental_tracer_with_nee(ctx):
    sc = N.gen_cornell(3)
    scene = N.Scene(ctx, sc)
    w, h = 40, 32
    p = N.RenderParams(w, h, 2, 8, 3, 1, 5, 0, 1)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    cam = N.camera_desc(sc["camera"])
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
    more = C.c_int32(1)
    while more.value:
        ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    ref = torch.zeros_like(img)
    st = scene.render(ref.data_ptr(), w, h, 2, 8, 3, 5, nee=1, mis=1)
    stats = N.RenderStats()
    ctx.check(N.lib().mrh_tracer_stats(t, C.byref(stats), None))
    assert stats.segments == st.segments
    # same paths; the per-pixel float additions may come in another order
    assert np.allclose(img.cpu().numpy(), ref.cpu().numpy(), rtol=1e-5, atol=1e-7)
    ctx.check(N.lib().mrh_tracer_destroy(t))
    scene.close()


def test_nee_errors(ctx):
    sc = N.gen_cornell(1)
    scene = N.Scene(ctx, sc)
    img = torch.zeros(16 * 16 * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(N.MrhError):
        scene.render(img.data_ptr(), 16, 16, 1, 8, 3, 0, nee=2)
    with pytest.raises(N.MrhError):
        scene.render(img.data_ptr(), 16, 16, 1, 8, 3, 0, nee=1, mis=3)
    p = N.RenderParams(16, 16, 1, 6, 3, 1, 0, 0, 2)
    g = N.GuideParams(2, 6, 0.5, 16)
    with pytest.raises(N.MrhError):                                   # direct_light_mis = 2
        ctx.check(N.lib().mrh_render_guided(ctx.h, scene.h, C.byref(N.camera_desc(sc["camera"])), C.byref(p),
                                            C.byref(g), None, img.data_ptr(), None))
    # a scene without lights renders black with NEE on
    dark = N.gen_sphere_grid(2, 6, 6, 1)
    ds = N.Scene(ctx, dark)
    ds.render(img.data_ptr(), 16, 16, 1, 6, 3, 0, nee=1, mis=1)
    assert float(img.abs().max()) == 0.0
    ds.close()
    scene.close()
