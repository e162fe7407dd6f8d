"""The drop-in boundary on the GPU: the C++ host (tests/cpp/tracer_system_host.cpp)
loads libmrh.so by path, resolves GenerateTracerSystem/DeleteTracerSystem and
drives GenerateWork -> while (Render()) -> Finalize; its image must equal the
one-call C-ABI render of the same frame bit for bit.  Also: the incremental
tracer's progressive accumulation, and the full C3 frame against the oracle."""
import os

import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")



def _render(ctx, sc, w, h, S, depth, rr, seed=0):
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), w, h, S, depth, rr, seed)
    out = img.cpu().numpy().reshape(h, w, 3)
    scene.close()
    return out, st


def test_incremental_tracer_matches_one_call_render(ctx):
    import ctypes as C
    sc = N.gen_cornell(3)
    scene = N.Scene(ctx, sc)
    w, h = 40, 32
    p = N.RenderParams(w, h, 2, 8, 3, 0, 5, 0)
    t = C.c_void_p()
    ctx.check(N.lib().mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    cam = N.camera_desc(sc["camera"])
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
    more = C.c_int32(1)
    n = 0
    while more.value:
        ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
        n += 1
    assert n == 8
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    one = img.cpu().numpy().reshape(h, w, 3)
    ref = torch.zeros_like(img)
    st = scene.render(ref.data_ptr(), w, h, 2, 8, 3, 5)
    assert one.tobytes() == ref.cpu().numpy().reshape(h, w, 3).tobytes()
    # a second frame continues the sample sequence and averages in
    ctx.check(N.lib().mrh_tracer_generate_work(t, C.byref(cam)))
    more.value = 1
    while more.value:
        ctx.check(N.lib().mrh_tracer_render(t, C.byref(more)))
    ctx.check(N.lib().mrh_tracer_finalize(t, img.data_ptr()))
    two = img.cpu().numpy().reshape(h, w, 3)
    stats = N.RenderStats()
    frames = C.c_uint32(0)
    ctx.check(N.lib().mrh_tracer_stats(t, C.byref(stats), C.byref(frames)))
    assert frames.value == 2 and stats.segments > st.segments
    assert not np.array_equal(one, two)
    assert abs(two.mean() - one.mean()) < 0.2 * one.mean()
    ctx.check(N.lib().mrh_tracer_destroy(t))
    scene.close()


def test_c3_full_frame_matches_oracle(ctx, oracle):
    """The bench frame (1920x1080, 16 spp, depth 10, 1,966,080 triangles):
    same segment count and relMSE <= 1e-4 against the oracle on host threads."""
    sc = N.gen_sphere_grid(64, 16, 16, 2)
    g, st = _render(ctx, sc, 1920, 1080, 4, 10, 3, 0)
    ob = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], 1920, 1080)
    threads = max(1, min(os.cpu_count() or 1, 16))
    ref, seg = oracle.render(sc, ob, cam, 1920, 1080, 4, 10, 3, 0, threads=threads)
    assert st.segments == seg
    rel = float(np.mean((g - ref) ** 2 / (ref ** 2 + 1e-2)))
    assert rel <= 1e-4, rel
