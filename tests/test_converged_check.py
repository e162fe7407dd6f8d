"""The converged-image check itself, on the CPU: the oracle rendered with
seeds other than the fixture's passes it, and the oracle with a planted
mistake (every Lambert albedo scaled by 0.9) fails it.  The GPU tests in
test_gpu_converged.py apply the same check to the HIP renders."""
import os

import numpy as np

from meturay_amd import native as N

import converged_stats as CS

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle_batches(oracle, sc, w, h, strata, depth, seed0, k=16):
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    imgs = [oracle.render(sc, bvh, cam, w, h, strata, depth, 3, seed0 + b, threads=os.cpu_count() or 4)[0]
            for b in range(k)]
    return CS.batch_mean_se(imgs)


def test_fixture_shape():
    f = np.load(os.path.join(HERE, "golden", "converged_c1_4096spp.npz"))
    assert f["mean"].shape == (256, 256, 3) and f["se"].shape == (256, 256, 3)
    assert int(f["spp"]) == 4096 and int(f["batches"]) == 16
    assert np.all(f["se"] >= 0) and np.all(np.isfinite(f["mean"]))


def test_check_accepts_oracle_and_rejects_planted_bias(oracle):
    f = np.load(os.path.join(HERE, "golden", "converged_grid8_4096spp.npz"))
    sc = N.gen_sphere_grid(8, 16, 16, 2)
    m, s = _oracle_batches(oracle, sc, 96, 64, 8, 12, 0x1234)
    r = CS.compare(m, s, f["mean"], f["se"], tile=4)
    assert not CS.failures(r), r
    bad = dict(sc)
    bad["lambert"] = (sc["lambert"] * 0.9).astype(np.float32)
    m, s = _oracle_batches(oracle, bad, 96, 64, 8, 12, 0x1234)
    r = CS.compare(m, s, f["mean"], f["se"], tile=4)
    assert CS.failures(r), r
