"""The serving engine: requests traced together must give the results they
give alone.

Each request's paths are keyed by its own seed and sample ids, so in the mix
it traces exactly the segments a solo mrh_render / mrh_render_guided call
traces, ends with the same guide, and its image differs from the solo image
only by the order of the per-pixel float additions.  Small queues force many
steps, so requests at different depths share every wavefront.
"""
import numpy as np
import pytest

from meturay_amd import native as N
import nee_scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (scene, width, height, strata, max depth, rr, seed, iterations, nee, mis)
MIX = [
    ("cornell", 48, 32, 2, 6, 2, 1, 1, 0, 0),
    ("spheres", 64, 48, 1, 8, 3, 2, 1, 0, 0),
    ("cornell", 40, 40, 2, 6, 2, 3, 3, 0, 0),
    ("lit", 32, 24, 2, 10, 3, 4, 3, 0, 0),
    ("cornell", 16, 16, 1, 4, 1, 5, 1, 1, 0),
    ("spheres", 80, 60, 1, 5, 2, 6, 2, 0, 0),
    ("lit", 48, 40, 1, 8, 3, 7, 1, 1, 1),
    ("cornell", 24, 24, 2, 6, 3, 8, 1, 1, 1),
    ("lit", 40, 30, 2, 8, 3, 9, 3, 1, 1),
]


@pytest.fixture(scope="module")
def scenes(ctx):
    sc = {"cornell": N.gen_cornell(2), "spheres": N.gen_sphere_grid(4, 8, 8, 5), "lit": S.lit_spheres()}
    res = {k: (v, N.Scene(ctx, v)) for k, v in sc.items()}
    yield res
    for _, s in res.values():
        s.close()


def solo(ctx, scene, w, h, spp_axis, depth, rr, seed, it, nee, mis):
    img = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    if it > 1:
        st = scene.render_guided(img.data_ptr(), w, h, spp_axis, depth, rr, seed, iterations=it, svo_depth=6,
                                 guide_ratio=0.5, min_samples=16, nee=nee, mis=mis)
    else:
        st = scene.render(img.data_ptr(), w, h, spp_axis, depth, rr, seed, nee=nee, mis=mis)
    return img.cpu().numpy(), st.segments


def submit_mix(server, scenes, imgs):
    tickets = {}
    for k, (name, w, h, spp_axis, depth, rr, seed, it, nee, mis) in enumerate(MIX):
        scene = scenes[name][1]
        tickets[server.submit(scene, imgs[k], w, h, spp_axis, depth, rr, seed, iterations=it, svo_depth=6,
                              guide_ratio=0.5, min_samples=16, tag=k, nee=nee, mis=mis)] = k
    return tickets


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("wave", [4096, 1 << 20])
def test_mix_matches_solo(ctx, scenes, graphs, wave):
    ref = [solo(ctx, scenes[m[0]][1], *m[1:]) for m in MIX]
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=wave, max_slots=4, use_graphs=graphs)
    imgs = [torch.full((m[1] * m[2] * 3,), 7.0, dtype=torch.float32, device="cuda") for m in MIX]
    torch.cuda.synchronize()
    tickets = submit_mix(server, scenes, imgs)
    done = server.run()
    assert sorted(d["tag"] for d in done) == list(range(len(MIX)))
    st = server.stats()
    assert st.active == 0 and st.queued == 0 and st.live_paths == 0
    if wave == 4096:
        assert st.steps > 20                                   # requests really shared steps
    for d in done:
        k = d["tag"]
        assert d["status"] == 0
        assert tickets[d["ticket"]] == k
        assert d["segments"] == ref[k][1], (k, d["segments"], ref[k][1])
        got = imgs[k].cpu().numpy()
        assert np.allclose(got, ref[k][0], rtol=1e-4, atol=1e-6), k
        assert d["submit_s"] <= d["admit_s"] <= d["done_s"]
    assert st.segments == sum(r[1] for r in ref)
    server.close()


def test_mix_matches_solo_global_partition(ctx, scenes):
    """The engine with the global radix partition instead of per-block tiles."""
    ref = [solo(ctx, scenes[m[0]][1], *m[1:]) for m in MIX]
    old = ctx.tune_get("shade_partition")
    ctx.tune("shade_partition", 0)
    try:
        server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=4, use_graphs=True)
    finally:
        ctx.tune("shade_partition", old)
    imgs = [torch.full((m[1] * m[2] * 3,), 7.0, dtype=torch.float32, device="cuda") for m in MIX]
    torch.cuda.synchronize()
    submit_mix(server, scenes, imgs)
    done = server.run()
    for d in done:
        k = d["tag"]
        assert d["status"] == 0 and d["segments"] == ref[k][1]
        assert np.allclose(imgs[k].cpu().numpy(), ref[k][0], rtol=1e-4, atol=1e-6), k
    server.close()


def test_budget_admission(ctx, scenes):
    # each plain 64x48 request holds its image (36 KiB): a 100 KiB budget
    # admits two at a time, the rest wait in the queue
    server = N.Server(ctx, hbm_budget=100 << 10, wave_paths=1 << 16, max_slots=8, use_graphs=True)
    scene = scenes["spheres"][1]
    imgs = [torch.zeros(64 * 48 * 3, dtype=torch.float32, device="cuda") for _ in range(5)]
    torch.cuda.synchronize()
    for k in range(5):
        server.submit(scene, imgs[k], 64, 48, 1, 6, 2, 10 + k, tag=k)
    peak = 0
    done = []
    while True:
        a, q = server.step()
        peak = max(peak, a)
        assert server.stats().budget_used <= 100 << 10
        done += server.poll()
        if a == 0 and q == 0:
            break
    assert peak == 2
    assert sorted(d["tag"] for d in done) == list(range(5))
    with pytest.raises(N.MrhError):
        server.submit(scene, imgs[0], 1024, 1024, 1, 6, 2, 0)     # needs 12 MiB > budget
    server.close()


def test_server_errors(ctx, scenes):
    with pytest.raises(N.MrhError):
        N.Server(ctx, wave_paths=100)
    with pytest.raises(N.MrhError):
        N.Server(ctx, max_slots=65)
    server = N.Server(ctx, hbm_budget=1 << 30, wave_paths=1 << 16, max_slots=4)
    scene = scenes["cornell"][1]
    img = torch.zeros(16 * 16 * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(N.MrhError):
        server.submit(scene, img, 16, 16, 1, 0, 1, 0)             # max_depth 0
    with pytest.raises(N.MrhError):
        server.submit(scene, img, 16, 16, 1, 4, 1, 0, iterations=2, svo_depth=11)
    with pytest.raises(N.MrhError):
        server.submit(scene, img, 16, 16, 1, 4, 1, 0, nee=2)
    assert server.step() == (0, 0)                                # idle step is a no-op
    server.close()
