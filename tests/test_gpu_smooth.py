"""Smooth shading on the GPU: scenes with per-corner normals render like
the oracle (segments equal, relMSE 1e-4) in the plain, NEE + MIS and guided
tracers; invalid normals and normals on a two-level scene are refused."""
import os

import numpy as np
import pytest

from meturay_amd import native as N
import nee_scenes as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.join(os.path.dirname(__file__), "golden", "scenes")


def relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


def _scene(name):
    if name == "spheres":
        sc = S.smooth_spheres()
        return sc, sc["camera"]
    sc = S.smooth_cornell()
    return sc, sc["camera"]


@pytest.mark.parametrize("nee", [0, 1])
@pytest.mark.parametrize("name", ["spheres", "cornell"])
def test_smooth_render_matches_oracle(ctx, oracle, name, nee):
    sc, c = _scene(name)
    assert sc["normals"].shape == (len(sc["tris"]), 9)
    w, h, S_, depth, rr, seed = 64, 48, 2, 8, 3, 11
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), w, h, S_, depth, rr, seed, camera=c, nee=nee, mis=nee)
    g = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    ob = oracle.LBVH(sc["tris"], 30)
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render(sc, ob, cam, w, h, S_, depth, rr, seed, threads=8,
                             tree_width=ctx.tune_get("trace_width"), nee=nee, mis=nee)
    assert st.segments == seg
    assert ref.mean() > 0.005
    assert relmse(g, ref) <= 1e-4
    # the normals matter: the flat render of the same scene differs
    flat = dict(sc)
    del flat["normals"]
    ref_flat, _ = oracle.render(flat, ob, cam, w, h, S_, depth, rr, seed, threads=8,
                                tree_width=ctx.tune_get("trace_width"), nee=nee, mis=nee)
    assert relmse(ref_flat, ref) > 1e-3


def test_smooth_guided_matches_oracle(ctx, oracle):
    from test_gpu_guide import assert_same_svo, gpu_view
    sc = S.smooth_spheres()
    w, h, S_, depth, rr, seed = 40, 30, 2, 6, 2, 3
    it, sd, ratio, ms = 3, 6, 0.5, 32
    scene = N.Scene(ctx, sc)
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    svo = N.Svo(ctx)
    torch.cuda.synchronize()
    st = scene.render_guided(img.data_ptr(), w, h, S_, depth, rr, seed, iterations=it, svo_depth=sd,
                             guide_ratio=ratio, min_samples=ms, svo=svo, nee=1, mis=1)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg, osvo = oracle.render_guided(sc, bvh, cam, w, h, S_, depth, rr, seed, it, sd, ratio, ms,
                                          threads=8, keep_svo=True, tree_width=ctx.tune_get("trace_width"),
                                          nee=1, mis=1)
    assert st.segments == seg
    assert_same_svo(gpu_view(ctx, svo), osvo)
    assert relmse(img.cpu().numpy().reshape(h, w, 3), ref) < 1e-4
    svo.close()
    osvo.close()
    scene.close()


def test_smooth_normals_refused(ctx):
    sc = S.smooth_spheres()
    bad = dict(sc)
    bad["normals"] = sc["normals"].copy()
    bad["normals"][7, 3:6] = 0.0                            # a zero normal
    with pytest.raises(N.MrhError) as e:
        N.Scene(ctx, bad)
    assert e.value.code == 1
    bad["normals"][7, 3:6] = np.nan
    with pytest.raises(N.MrhError) as e:
        N.Scene(ctx, bad)
    assert e.value.code == 1
    two = N.gen_instanced_sphere_grid(4, 8, 8, 5, meshes=4)
    two["normals"] = S.smooth_normals(two["tris"])
    with pytest.raises(N.MrhError) as e:
        N.Scene(ctx, two)
    assert e.value.code == 4
