"""Converged images: HIP renders from seeds other than the fixture's,
compared with the oracle's 4096-spp FP64-accumulated images
(tests/golden/converged_*_4096spp.npz, written by gen_converged.py) by the
statistical check of converged_stats.py (per-pixel and per-tile z-scores
against the two sides' standard errors, relMSE against its expectation).

Plain path tracing, next-event estimation with MIS, the guided tracer
(4 learn-then-render iterations, octree depth 10, 50/50 mixture, 512-sample
node threshold) and the field-traced guided tracer (mrh_render_wfpg: voxel
octree depth 10, bins of 512 rays from level 5, per-bin cone-traced fields,
50/50 mixture; and its product mode, the guide's half sampling field x
BxDF) must all converge to the plain oracle's image: each is an unbiased estimator of
the same radiance."""
import os

import numpy as np
import pytest

from meturay_amd import native as N

import converged_stats as CS

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
SEED0 = 0x7E570000
BATCHES = 16

SCENES = {
    # name: (generator, w, h, depth, tile, strata per batch)
    "c1": (lambda: N.gen_cornell(4), 256, 256, 8, 8, 8),
    "grid8": (lambda: N.gen_sphere_grid(8, 16, 16, 2), 96, 64, 12, 4, 16),
}


def _gpu_batches(ctx, name, mode):
    gen, w, h, depth, tile, strata = SCENES[name]
    sc = gen()
    scene = N.Scene(ctx, sc)
    img = torch.zeros(h * w * 3, dtype=torch.float32, device="cuda")
    out = []
    for b in range(BATCHES):
        torch.cuda.synchronize()
        if mode == "plain":
            scene.render(img.data_ptr(), w, h, strata, depth, 3, SEED0 + b)
        elif mode == "nee_mis":
            scene.render(img.data_ptr(), w, h, strata, depth, 3, SEED0 + b, nee=1, mis=1)
        elif mode == "wfpg":   # the field-traced guide, same iteration split
            _, ws = scene.render_wfpg(img.data_ptr(), w, h, strata // 2, depth, 3, SEED0 + b, iterations=4,
                                      svo_depth=10, guide_ratio=0.5, min_bin_level=5, bin_ray_count=512)
            assert ws.guided_bounces > 0 and ws.bins > 0 and ws.binned_rays > 0
        elif mode == "wfpg_product":   # the same, the guide sampling field x BxDF
            _, ws = scene.render_wfpg(img.data_ptr(), w, h, strata // 2, depth, 3, SEED0 + b, iterations=4,
                                      svo_depth=10, guide_ratio=0.5, min_bin_level=5, bin_ray_count=512,
                                      product=1)
            assert ws.guided_bounces > 0 and ws.bins > 0 and ws.binned_rays > 0
        else:   # guided: 4 iterations of (strata/2)^2 spp = the same spp per batch
            scene.render_guided(img.data_ptr(), w, h, strata // 2, depth, 3, SEED0 + b, iterations=4,
                                svo_depth=10, guide_ratio=0.5, min_samples=512)
        out.append(img.cpu().numpy().reshape(h, w, 3).astype(np.float64))
    scene.close()
    return CS.batch_mean_se(out), tile


@pytest.mark.parametrize("mode", ["plain", "nee_mis", "guided", "wfpg", "wfpg_product"])
@pytest.mark.parametrize("name", ["c1", "grid8"])
def test_converges_to_fp64_oracle(ctx, name, mode):
    print(f"[converged] {name} {mode}", flush=True)
    f = np.load(os.path.join(HERE, "golden", f"converged_{name}_4096spp.npz"))
    (m, s), tile = _gpu_batches(ctx, name, mode)
    r = CS.compare(m, s, f["mean"], f["se"], tile=tile)
    print(r, flush=True)
    assert not CS.failures(r), (CS.failures(r), r)
