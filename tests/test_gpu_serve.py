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
    ("smooth", 48, 36, 1, 6, 2, 10, 1, 1, 1),
    ("smooth", 32, 32, 2, 6, 2, 11, 2, 0, 0),
]


@pytest.fixture(scope="module")
def scenes(ctx):
    sc = {"cornell": N.gen_cornell(2), "spheres": N.gen_sphere_grid(4, 8, 8, 5), "lit": S.lit_spheres(),
          "smooth": S.smooth_spheres()}
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


def _fixed_bytes(ctx, wave):
    probe = N.Server(ctx, hbm_budget=8 << 30, wave_paths=wave, max_slots=8)
    fixed = probe.stats().fixed_bytes
    probe.close()
    return fixed


def test_budget_admission(ctx, scenes):
    """The budget covers the queues, each admitted request's scene (once) and
    its image: beside the queues and the scene, 100 KiB admits two plain
    64x48 requests (36 KiB images) at a time, the rest wait in the queue."""
    wave = 1 << 16
    fixed = _fixed_bytes(ctx, wave)
    scene = scenes["spheres"][1]
    sb = scene.bytes()
    budget = fixed + sb + (100 << 10)
    server = N.Server(ctx, hbm_budget=budget, wave_paths=wave, max_slots=8, use_graphs=True)
    st = server.stats()
    assert st.fixed_bytes == fixed and st.budget_used == fixed and st.scene_bytes == 0
    imgs = [torch.zeros(64 * 48 * 3, dtype=torch.float32, device="cuda") for _ in range(5)]
    torch.cuda.synchronize()
    for k in range(5):
        server.submit(scene, imgs[k], 64, 48, 1, 6, 2, 10 + k, tag=k)
    peak = 0
    done = []
    while True:
        a, q = server.step()
        peak = max(peak, a)
        st = server.stats()
        assert st.budget_used <= budget
        if a:
            assert st.scene_bytes == sb and st.budget_used == fixed + sb + a * 64 * 48 * 12
        done += server.poll()
        if a == 0 and q == 0:
            break
    assert peak == 2
    assert sorted(d["tag"] for d in done) == list(range(5))
    st = server.stats()
    assert st.budget_used == fixed and st.scene_bytes == 0        # the scene is released with its last request
    with pytest.raises(N.MrhError):
        server.submit(scene, imgs[0], 1024, 1024, 1, 6, 2, 0)     # needs 12 MiB > budget
    server.close()
    with pytest.raises(N.MrhError):
        N.Server(ctx, hbm_budget=fixed - 1, wave_paths=wave)      # the queues alone do not fit


def test_scene_charged_once(ctx, scenes):
    """Two admitted requests on one scene charge it once; a request on a
    second scene waits until its scene fits beside the first."""
    wave = 1 << 16
    fixed = _fixed_bytes(ctx, wave)
    a, b = scenes["spheres"][1], scenes["cornell"][1]
    img = 32 * 32 * 12
    budget = fixed + max(a.bytes(), b.bytes()) + 3 * img
    server = N.Server(ctx, hbm_budget=budget, wave_paths=wave, max_slots=8, use_graphs=False)
    imgs = [torch.zeros(32 * 32 * 3, dtype=torch.float32, device="cuda") for _ in range(3)]
    torch.cuda.synchronize()
    server.submit(a, imgs[0], 32, 32, 2, 6, 2, 1, tag=0)
    server.submit(a, imgs[1], 32, 32, 2, 6, 2, 2, tag=1)
    server.submit(b, imgs[2], 32, 32, 2, 6, 2, 3, tag=2)
    act, q = server.step()
    assert (act, q) == (2, 1)                                     # b's scene does not fit beside a's
    assert server.stats().scene_bytes == a.bytes()
    done = server.run()
    assert sorted(d["tag"] for d in done) == [0, 1, 2] and all(d["status"] == 0 for d in done)
    server.close()


def test_cancel_and_timeout(ctx, scenes):
    """A cancelled request ends alone with MRH_ERR_CANCELLED, queued or in
    flight; a request older than timeout_ms ends with MRH_ERR_TIMEOUT; the
    others still equal their solo renders."""
    CANCELLED, TIMEOUT = 6, 7
    jobs = [("cornell", 48, 32, 2, 6, 2, 11, 1, 0, 0), ("spheres", 64, 48, 2, 8, 3, 12, 2, 0, 0),
            ("lit", 32, 24, 2, 10, 3, 13, 1, 1, 1), ("cornell", 40, 40, 2, 6, 2, 14, 1, 0, 0)]
    ref = [solo(ctx, scenes[m[0]][1], *m[1:]) for m in jobs]
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=2, use_graphs=True)
    imgs = [torch.zeros(m[1] * m[2] * 3, dtype=torch.float32, device="cuda") for m in jobs]
    torch.cuda.synchronize()
    t = []
    for k, (name, w, h, sa, depth, rr, seed, it, nee, mis) in enumerate(jobs):
        t.append(server.submit(scenes[name][1], imgs[k], w, h, sa, depth, rr, seed, iterations=it, svo_depth=6,
                               guide_ratio=0.5, min_samples=16, tag=k, nee=nee, mis=mis))
    server.step()
    server.step()
    server.cancel(t[1])                  # admitted, paths in flight, learning
    server.cancel(t[3])                  # still queued: reported at once
    with pytest.raises(N.MrhError):
        server.cancel(12345)
    done = server.run()
    by = {d["tag"]: d for d in done}
    assert sorted(by) == [0, 1, 2, 3]
    assert by[1]["status"] == CANCELLED and by[3]["status"] == CANCELLED
    for k in (0, 2):
        assert by[k]["status"] == 0 and by[k]["segments"] == ref[k][1]
        assert np.allclose(imgs[k].cpu().numpy(), ref[k][0], rtol=1e-4, atol=1e-6), k
    st = server.stats()
    assert st.active == 0 and st.live_paths == 0 and st.budget_used == st.fixed_bytes
    with pytest.raises(N.MrhError):
        server.cancel(t[0])              # already done
    server.close()
    # a 1 ms limit: every request times out, none blocks the server
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=2, use_graphs=True, timeout_ms=1)
    for k, (name, w, h, sa, depth, rr, seed, it, nee, mis) in enumerate(jobs):
        server.submit(scenes[name][1], imgs[k], w, h, sa, depth, rr, seed, iterations=it, svo_depth=6,
                      guide_ratio=0.5, min_samples=16, tag=k, nee=nee, mis=mis)
    import time
    time.sleep(0.01)
    done = server.run()
    assert sorted(d["tag"] for d in done) == [0, 1, 2, 3]
    assert all(d["status"] == TIMEOUT for d in done)
    assert server.stats().budget_used == server.stats().fixed_bytes
    server.close()


def test_injected_allocation_failure_is_isolated(ctx, scenes):
    """The guided request's learn-record allocation fails (injected on the
    server's context): that request alone ends with MRH_ERR_OOM, the step
    still succeeds, and the plain requests equal their solo renders."""
    OOM = 3
    jobs = [("cornell", 48, 32, 2, 6, 2, 21, 1, 0, 0), ("spheres", 40, 30, 2, 8, 3, 22, 3, 0, 0),
            ("lit", 32, 24, 2, 10, 3, 23, 1, 1, 1)]
    ref = [solo(ctx, scenes[m[0]][1], *m[1:]) for m in jobs]
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=4, use_graphs=True)
    imgs = [torch.zeros(m[1] * m[2] * 3, dtype=torch.float32, device="cuda") for m in jobs]
    torch.cuda.synchronize()
    for k, (name, w, h, sa, depth, rr, seed, it, nee, mis) in enumerate(jobs):
        server.submit(scenes[name][1], imgs[k], w, h, sa, depth, rr, seed, iterations=it, svo_depth=6,
                      guide_ratio=0.5, min_samples=16, tag=k, nee=nee, mis=mis)
    # admission allocates nothing for plain requests: the first allocation
    # after this point is the guided request's record buffer
    server.inner.tune("fail_alloc_after", 1)
    done = server.run()
    assert server.inner.tune_get("fail_alloc_after") == 0
    by = {d["tag"]: d for d in done}
    assert sorted(by) == [0, 1, 2]
    assert by[1]["status"] == OOM
    for k in (0, 2):
        assert by[k]["status"] == 0 and by[k]["segments"] == ref[k][1]
        assert np.allclose(imgs[k].cpu().numpy(), ref[k][0], rtol=1e-4, atol=1e-6), k
    st = server.stats()
    assert st.active == 0 and st.budget_used == st.fixed_bytes
    server.close()


def test_mix_matches_solo_in_arrival_order(ctx, scenes):
    """Injection in submission order (policy MRH_SERVE_ARRIVAL) changes only
    which request's samples enter a step first, never their paths."""
    ref = [solo(ctx, scenes[m[0]][1], *m[1:]) for m in MIX]
    server = N.Server(ctx, hbm_budget=8 << 30, wave_paths=4096, max_slots=4, use_graphs=True, policy=1)
    imgs = [torch.full((m[1] * m[2] * 3,), 7.0, dtype=torch.float32, device="cuda") for m in MIX]
    torch.cuda.synchronize()
    submit_mix(server, scenes, imgs)
    done = server.run()
    assert sorted(d["tag"] for d in done) == list(range(len(MIX)))
    for d in done:
        k = d["tag"]
        assert d["status"] == 0 and d["segments"] == ref[k][1]
        assert np.allclose(imgs[k].cpu().numpy(), ref[k][0], rtol=1e-4, atol=1e-6), k
    server.close()


def test_server_errors(ctx, scenes):
    with pytest.raises(N.MrhError):
        N.Server(ctx, policy=2)
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
