"""The drop-in for TracerCUDA on the GPU: tests/cpp/mray_tracer_thread loads
libTracerMRH.so by path, resolves GenerateTracerSystem / DeleteTracerSystem
and runs MRay's TracerThread sequence through the reference's own
interfaces (Initialize, GenerateScene(path) + LoadScene, GenerateTracer,
GenerateReconFilter + AttachReconFilter, ResizeImage, ReportionImage,
GenerateWork -> while (Render()) -> Finalize).  The image the visor
accumulates from SendImage must equal the C-ABI render of the same scene
file, camera, seed and filter."""
import os
import subprocess

import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "meturay_amd", "adapter", "_build", "libTracerMRH.so")
HOST = os.path.join(ROOT, "tests", "cpp", "_build", "mray_tracer_thread")
SCENE = os.path.join(ROOT, "tests", "golden", "scenes", "room.json")
W, H = 64, 48


def _host(tmp_path, tracer, frames, samples, seed, filt):
    assert os.path.exists(ADAPTER) and os.path.exists(HOST), "adapter not built (make -C meturay_amd/adapter all host)"
    out = tmp_path / "img.bin"
    r = subprocess.run([HOST, "gpu", ADAPTER, SCENE, tracer, str(W), str(H), str(frames), str(samples), str(seed),
                        filt, str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    return np.fromfile(out, np.float32).reshape(H, W, 3)


@pytest.mark.parametrize("filt", ["box0", "tent1"])
def test_path_tracer_through_mray_interfaces(ctx, tmp_path, filt):
    got = _host(tmp_path, "PathTracer", 1, 2, 5, filt)
    sc = N.load_scene_file(SCENE)
    scene = N.Scene(ctx, sc)
    img = torch.zeros(W * H * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    f = N.filter_desc("Box", 0.0) if filt == "box0" else N.filter_desc("Tent", 1.0)
    scene.render_filtered(img.data_ptr(), W, H, 2, 8, 3, 5, filt=f, camera=sc["cameras"][0])
    ref = img.cpu().numpy().reshape(H, W, 3)
    scene.close()
    assert ref.mean() > 0.005
    # one frame of 4 spp: value x 4 / 4 is exact, so the visor's image is the render's
    assert got.tobytes() == ref.tobytes()


def _incremental(ctx, scene, cam, frames, samples, seed, guide=None):
    """frames of the C-ABI incremental tracer, Box radius 0 filter; guide =
    mrh_wfpg_params fields (iterations, octree depth, ratio, min bin level,
    bin ray count, filter alpha, boundary radiance, product) or None"""
    import ctypes as C
    L = N.lib()
    p = N.RenderParams(W, H, samples, 8, 3, 0, seed, 0, 0)
    t = C.c_void_p()
    ctx.check(L.mrh_tracer_create(ctx.h, scene.h, C.byref(p), C.byref(t)))
    f = N.filter_desc("Box", 0.0)
    ctx.check(L.mrh_tracer_set_filter(t, C.byref(f)))
    if guide:
        g = N.WfpgParams(*guide)
        ctx.check(L.mrh_tracer_set_wfpg(t, C.byref(g)))
    cd = N.camera_desc(cam)
    more = C.c_int32(1)
    for _ in range(frames):
        ctx.check(L.mrh_tracer_generate_work(t, C.byref(cd)))
        more.value = 1
        while more.value:
            ctx.check(L.mrh_tracer_render(t, C.byref(more)))
    img = torch.zeros(W * H * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.check(L.mrh_tracer_finalize(t, img.data_ptr()))
    ctx.check(L.mrh_tracer_destroy(t))
    return img.cpu().numpy().reshape(H, W, 3)


def test_frames_accumulate_and_guided_tracer(ctx, tmp_path):
    sc = N.load_scene_file(SCENE)
    scene = N.Scene(ctx, sc)
    cam = sc["cameras"][0]
    # three frames: the visor's running mean is the tracer's progressive mean
    got = _host(tmp_path, "PathTracer", 3, 1, 9, "box0")
    ref = _incremental(ctx, scene, cam, 3, 1, 9)
    assert ref.mean() > 0.005
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-7)
    # WFPGTracer (octree level 8, bins of 64 rays from level 4, ratio 0.5,
    # alpha 1, no product): the field-traced guide, its payload learning over
    # the first 16 frames, the vertices of frame k guiding frame k + 1
    got = _host(tmp_path, "WFPGTracer", 3, 1, 9, "box0")
    ref = _incremental(ctx, scene, cam, 3, 1, 9, guide=(16, 8, 0.5, 4, 64, 1.0, 0.0, 0))
    scene.close()
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-7)
