"""Reconstruction filters on the GPU, through the C-ABI, against the oracle.

The gather sums every pixel's contributions in the oracle's order, so Box,
Tent and Mitchell results are bit-identical; the Gaussian differs only in
the last bits of the device expf (tolerance 1e-6 relative).  Filtered
renders without MIS keep every sample's radiance exactly, so they are
bit-identical too (with MIS two contributions of one sample may arrive in
either order).
"""
import ctypes as C
import os

import numpy as np
import pytest

from meturay_amd import native as N
import nee_scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")



def descs(kind, r):
    if kind == "Gaussian":
        return N.filter_desc(kind, r, alpha=2.0)
    if kind == "Mitchell":
        return N.filter_desc(kind, r, b=1.0 / 3.0, c=1.0 / 3.0)
    return N.filter_desc(kind, r)


def odesc(oracle, d):
    return oracle.OmFilterDesc(d.type, d.radius, d.alpha, d.b, d.c)


def same(kind, got, ref):
    if kind == "Gaussian":
        return np.allclose(got, ref, rtol=1e-6, atol=1e-7)
    return np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["Box", "Tent", "Gaussian", "Mitchell"])
@pytest.mark.parametrize("r", [0.0, 0.5, 1.0, 1.5, 2.5])
def test_filter_samples_matches_oracle(ctx, oracle, kind, r):
    if r == 0.0 and kind != "Box":
        pytest.skip("radius 0 is the Box filter's")
    rng = np.random.default_rng(int(10 * r) + 7)
    w, h, n = 64, 48, 60000
    coords = np.stack([rng.uniform(-1, w + 1, n), rng.uniform(-1, h + 1, n)], 1).astype(np.float32)
    coords[:500] = np.floor(coords[:500]) + 0.5
    values = rng.uniform(0, 3, (n, 3)).astype(np.float32)
    d = descs(kind, r)
    wtd = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    wts = torch.zeros(h * w, dtype=torch.float32, device="cuda")
    dv = torch.from_numpy(values).cuda()
    dc = torch.from_numpy(coords).cuda()
    torch.cuda.synchronize()
    ctx.filter_samples(d, dv, dc, n, w, h, wtd, wts)
    rw, rws = oracle.filter_samples(odesc(oracle, d), values, coords, w, h)
    assert same(kind, wtd.cpu().numpy().reshape(h, w, 3), rw)
    assert same(kind, wts.cpu().numpy().reshape(h, w), rws)
    # a second call accumulates; resolve divides
    ctx.filter_samples(d, dv, dc, n, w, h, wtd, wts)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    ctx.filter_resolve(wtd, wts, w, h, img)
    ref = oracle.filter_resolve(rw + rw, rws + rws)
    assert same(kind, img.cpu().numpy().reshape(h, w, 3), ref)


def test_filter_errors(ctx):
    wtd = torch.zeros(16 * 3, dtype=torch.float32, device="cuda")
    wts = torch.zeros(16, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for bad in [N.FilterDesc(7, 1.0, 0, 0, 0), N.filter_desc("Tent", 0.0), N.filter_desc("Box", -1.0),
                N.filter_desc("Gaussian", 1.0, alpha=0.0)]:
        with pytest.raises(N.MrhError):
            ctx.filter_samples(bad, wtd, wtd, 0, 4, 4, wtd, wts)
    ctx.filter_samples(N.filter_desc("Tent", 1.0), None, None, 0, 4, 4, wtd, wts)   # no samples: a no-op
    assert float(wts.abs().sum()) == 0.0


def _filtered_pair(ctx, oracle, sc, w, h, spp_axis, depth, rr, seed, d, nee=0, mis=0):
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render_filtered(img.data_ptr(), w, h, spp_axis, depth, rr, seed, filt=d, nee=nee, mis=mis)
    g = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    ob = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    ref, seg = oracle.render_filtered(sc, ob, cam, w, h, spp_axis, depth, rr, seed, odesc(oracle, d), threads=8,
                                      tree_width=ctx.tune_get("trace_width"), nee=nee, mis=mis)
    return g, st, ref, seg


@pytest.mark.parametrize("kind,r", [("Box", 0.0), ("Tent", 1.0), ("Mitchell", 2.0), ("Gaussian", 1.5)])
def test_render_filtered_matches_oracle(ctx, oracle, kind, r):
    d = descs(kind, r)
    g, st, ref, seg = _filtered_pair(ctx, oracle, N.gen_cornell(3), 64, 48, 2, 8, 3, 3, d)
    assert st.segments == seg
    assert ref.mean() > 0.01
    assert same(kind, g, ref)


def test_render_filtered_with_nee(ctx, oracle):
    d = descs("Tent", 1.0)
    g, st, ref, seg = _filtered_pair(ctx, oracle, S.lit_spheres(), 96, 64, 2, 8, 3, 4, d, nee=1, mis=0)
    assert st.segments == seg and np.array_equal(g, ref)
    g, st, ref, seg = _filtered_pair(ctx, oracle, S.lit_spheres(), 96, 64, 2, 8, 3, 4, d, nee=1, mis=1)
    assert st.segments == seg and np.allclose(g, ref, rtol=1e-5, atol=1e-6)


def test_box_zero_equals_render(ctx):
    sc = N.gen_cornell(3)
    scene = N.Scene(ctx, sc)
    a = torch.zeros(40 * 32 * 3, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    torch.cuda.synchronize()
    scene.render(a.data_ptr(), 40, 32, 3, 8, 3, 9)
    scene.render_filtered(b.data_ptr(), 40, 32, 3, 8, 3, 9, filt=N.filter_desc("Box", 0.0))
    assert np.allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-6, atol=1e-7)
    scene.close()


def test_tracer_filter_accumulates_frames(ctx, oracle):
    sc = N.gen_cornell(3)
    scene = N.Scene(ctx, sc)
    w, h, spp_axis = 32, 24, 2
    p = N.RenderParams(w, h, spp_axis, 6, 3, 0, 2, 0, 0)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    d = descs("Mitchell", 2.0)
    ctx.check(N.lib().mrh_tracer_set_filter(t, C.byref(d)))
    cam = N.camera_desc(sc["camera"])
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(2):
        ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
        more = C.c_int32(1)
        while more.value:
            ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    got = img.cpu().numpy().reshape(h, w, 3)
    # oracle: frame k's sample ids start at k * W * H * spp; the tracer adds
    # both frames' filter sums before dividing
    ob = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    ocam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    od = odesc(oracle, d)
    tw = ctx.tune_get("trace_width")
    total = w * h * spp_axis * spp_axis
    sums = []
    for k in range(2):
        v, co, _ = oracle.render_samples(sc, ob, ocam, w, h, spp_axis, 6, 3, 2, threads=8, tree_width=tw,
                                         sample_base=k * total)
        sums.append(oracle.filter_samples(od, v, co, w, h, spp_axis * spp_axis))
    ref = oracle.filter_resolve(sums[0][0] + sums[1][0], sums[0][1] + sums[1][1])
    assert np.array_equal(got, ref)
    # reset starts the accumulation over
    ctx.check(N.lib().mrh_tracer_reset_image(t))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    assert float(img.abs().max()) == 0.0
    # detaching the filter goes back to the per-pixel mean
    ctx.check(N.lib().mrh_tracer_set_filter(t, None))
    ctx.check(N.lib().mrh_tracer_destroy(t))
    scene.close()
