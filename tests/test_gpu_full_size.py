"""The guided and serving workloads at their full configurations.

Guided (C4): the 64x64 displaced-sphere grid (1,966,080 triangles, seed 2),
1920x1080, 4 learn-then-render iterations of 16 spp, max depth 12, RR after
3, octree depth 10, 50/50 guide/BxDF mixture, 512-sample node threshold.
The HIP render must trace exactly the oracle's segments, end with the
oracle's octree bit for bit (about 2 M nodes: links, counts, 48.16 fixed-point
sums, filtered radiance), and give the oracle's image within relMSE 1e-4.

Serving (C5): the first 40 requests of the serving stream (2000 requests,
Poisson 40/s, seed 3) plus the stream's first 1920x1080 64-spp guided
request, over its 8 resident sphere-grid scenes (18K to 1.2M triangles),
through the engine with the benchmark's settings (200 GB budget, 2^24 paths
per wave, 64 slots, hipGraph steps).  Every request must trace the segments
it traces alone and give its solo image (float-add order aside); one plain
and one guided request are also checked against the oracle.
"""
import os

import numpy as np
import pytest

from meturay_amd import native as N
from meturay_amd.workloads import serving_scene, serving_workload

from test_gpu_guide import assert_same_svo, gpu_view

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

THREADS = max(1, min(os.cpu_count() or 1, 16))


def _relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


def _oracle_cam(oracle, sc, w, h):
    c = sc["camera"]
    return oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)


@pytest.mark.timeout(900)
def test_guided_full_size_matches_oracle(ctx, oracle):
    W, H, S, depth, rr, seed, it, svo_depth, ratio, ms = 1920, 1080, 4, 12, 3, 1, 4, 10, 0.5, 512
    sc = N.gen_sphere_grid(64, 16, 16, 2)
    scene = N.Scene(ctx, sc)
    img = torch.zeros(W * H * 3, dtype=torch.float32, device="cuda")
    svo = N.Svo(ctx)
    torch.cuda.synchronize()
    st = scene.render_guided(img.data_ptr(), W, H, S, depth, rr, seed, iterations=it, svo_depth=svo_depth,
                             guide_ratio=ratio, min_samples=ms, svo=svo)
    got = img.cpu().numpy().reshape(H, W, 3)
    g = gpu_view(ctx, svo)
    print(f"[full-size] gpu: {st.segments} segments, octree {g['nodes']} nodes; oracle ...", flush=True)
    bvh = oracle.LBVH(sc["tris"], 30)
    ref, seg, osvo = oracle.render_guided(sc, bvh, _oracle_cam(oracle, sc, W, H), W, H, S, depth, rr, seed,
                                          it, svo_depth, ratio, ms, threads=THREADS, keep_svo=True,
                                          tree_width=ctx.tune_get("trace_width"))
    print(f"[full-size] oracle: {seg} segments, octree {osvo.nodes} nodes", flush=True)
    assert st.segments == seg
    assert g["nodes"] > 1_000_000
    assert_same_svo(g, osvo)
    err = _relmse(got, ref)
    assert err <= 1e-4, err
    svo.close()
    osvo.close()
    scene.close()


def _requests():
    grids, reqs = serving_workload(2000, 40.0)
    pick = list(range(40))
    big = next(i for i, r in enumerate(reqs)
               if (r["w"], r["h"], r["spp"]) == (1920, 1080, 64) and r["it"] > 1)
    pick.append(big)
    return grids, [reqs[i] for i in pick]


@pytest.mark.timeout(900)
def test_serving_full_size_mix_equals_solo(ctx, oracle):
    depth, rr = 10, 3
    grids, reqs = _requests()
    scs = [serving_scene(N, g, k) for k, g in enumerate(grids)]
    scenes = [N.Scene(ctx, sc) for sc in scs]
    srv = N.Server(ctx, hbm_budget=200 << 30, wave_paths=1 << 24, max_slots=64, use_graphs=True)
    imgs = [torch.full((r["w"] * r["h"] * 3,), 7.0, dtype=torch.float32, device="cuda") for r in reqs]
    torch.cuda.synchronize()
    for k, r in enumerate(reqs):
        srv.submit(scenes[r["scene"]], imgs[k], r["w"], r["h"], r["S"], depth, rr, r["seed"],
                   iterations=r["it"], svo_depth=10, guide_ratio=0.5, min_samples=512, tag=k)
    done = srv.run()
    st = srv.stats()
    srv.close()
    print(f"[full-size] engine: {len(done)} requests, {st.steps} steps, {st.segments} segments", flush=True)
    assert sorted(d["tag"] for d in done) == list(range(len(reqs)))
    seg_mix = {d["tag"]: d["segments"] for d in done}
    assert all(d["status"] == 0 for d in done)
    solo = torch.empty(1920 * 1080 * 3, dtype=torch.float32, device="cuda")
    for k, r in enumerate(reqs):
        sc = scenes[r["scene"]]
        n = r["w"] * r["h"] * 3
        torch.cuda.synchronize()
        if r["it"] > 1:
            s = sc.render_guided(solo.data_ptr(), r["w"], r["h"], r["S"], depth, rr, r["seed"], iterations=r["it"],
                                 svo_depth=10, guide_ratio=0.5, min_samples=512)
        else:
            s = sc.render(solo.data_ptr(), r["w"], r["h"], r["S"], depth, rr, r["seed"])
        assert seg_mix[k] == s.segments, (k, r, seg_mix[k], s.segments)
        a = imgs[k].cpu().numpy()
        b = solo[:n].cpu().numpy()
        assert np.allclose(a, b, rtol=1e-4, atol=1e-6), (k, r)
    # one plain and one guided request against the oracle
    plain = next(k for k, r in enumerate(reqs) if r["it"] == 1 and r["w"] == 256)
    guided = next(k for k, r in enumerate(reqs) if r["it"] > 1 and r["w"] == 256)
    for k in (plain, guided):
        r = reqs[k]
        sc = scs[r["scene"]]
        bvh = oracle.LBVH(sc["tris"], 30)
        cam = _oracle_cam(oracle, sc, r["w"], r["h"])
        tw = ctx.tune_get("trace_width")
        if r["it"] > 1:
            ref, seg, _ = oracle.render_guided(sc, bvh, cam, r["w"], r["h"], r["S"], depth, rr, r["seed"], r["it"],
                                               10, 0.5, 512, threads=THREADS, tree_width=tw)
        else:
            ref, seg = oracle.render(sc, bvh, cam, r["w"], r["h"], r["S"], depth, rr, r["seed"], threads=THREADS,
                                     tree_width=tw)
        assert seg_mix[k] == seg, (k, r)
        got = imgs[k].cpu().numpy().reshape(r["h"], r["w"], 3)
        assert _relmse(got, ref) <= 1e-4, (k, r)
    for s in scenes:
        s.close()
