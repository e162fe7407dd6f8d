"""Launch-shape knobs and the autotuner on the GPU: every traversal block
size and partition tile renders the frame the default shape renders (same
segments, same image up to the order of float additions), and mrh_autotune
leaves a valid choice on the context, scored by a true median of at least
three renders per candidate."""
import itertools

import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

W, H, S, DEPTH, RR, SEED = 96, 64, 2, 8, 3, 5
SETS = {"trace_width": {2, 4}, "trace_lds_entries": {0, 8, 12, 16, 24}, "trace_blocks_per_cu": {4, 8, 16, 32},
        "trace_block": {128, 256, 512}, "shade_tile": {512, 1024, 2048}}


def render(scene, img, **kw):
    img.zero_()
    torch.cuda.synchronize()
    st = scene.render(img.data_ptr(), W, H, S, DEPTH, RR, SEED, **kw)
    return st.segments, img.cpu().numpy().copy()


@pytest.fixture
def knobs(ctx):
    old = {k: ctx.tune_get(k) for k in SETS}
    old["shade_partition"] = ctx.tune_get("shade_partition")
    yield
    for k, v in old.items():
        ctx.tune(k, v)


@pytest.mark.parametrize("nee", [0, 1])
def test_every_shape_renders_the_same_frame(ctx, knobs, nee):
    scene = N.Scene(ctx, N.gen_cornell(2))
    img = torch.zeros(W * H * 3, dtype=torch.float32, device="cuda")
    seg0, ref = render(scene, img, nee=nee, mis=nee)
    for width, lds, block, tile in itertools.product([2, 4], [0, 8, 24], [128, 256, 512], [512, 1024, 2048]):
        ctx.tune("trace_width", width)
        ctx.tune("trace_lds_entries", lds)
        ctx.tune("trace_block", block)
        ctx.tune("shade_tile", tile)
        seg, got = render(scene, img, nee=nee, mis=nee)
        assert seg == seg0, (width, lds, block, tile)
        assert np.allclose(got, ref, rtol=1e-5, atol=1e-6), (width, lds, block, tile)
    # the global partition with a non-default traversal block
    ctx.tune("shade_partition", 0)
    seg, got = render(scene, img, nee=nee, mis=nee)
    assert seg == seg0 and np.allclose(got, ref, rtol=1e-5, atol=1e-6)
    scene.close()


def test_knob_limits(ctx, knobs):
    for k, bad in [("trace_block", 64), ("trace_block", 1024), ("shade_tile", 256), ("shade_tile", 4096)]:
        with pytest.raises(N.MrhError):
            ctx.tune(k, bad)
    ctx.tune("trace_block", 512)
    ctx.tune("shade_tile", 2048)
    assert ctx.tune_get("trace_block") == 512 and ctx.tune_get("shade_tile") == 2048


def test_autotune_leaves_a_valid_choice(ctx, knobs):
    scene = N.Scene(ctx, N.gen_cornell(2))
    img = torch.zeros(W * H * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    res = scene.autotune(img.data_ptr(), W, H, S, DEPTH, RR, SEED, reps=1)
    assert res.reps == 3                                   # at least three renders per candidate
    assert res.trials == sum(len(v) for v in SETS.values())
    best = {"trace_width": res.best_width, "trace_lds_entries": res.best_lds, "trace_blocks_per_cu": res.best_bpc,
            "trace_block": res.best_block, "shade_tile": res.best_tile}
    for k, v in best.items():
        assert v in SETS[k], k
        assert ctx.tune_get(k) == v, k                     # left set on the context
    trials = [(res.trial_width[i], res.trial_lds[i], res.trial_bpc[i], res.trial_block[i], res.trial_tile[i],
               res.trial_ms[i]) for i in range(res.trials)]
    for t in trials:
        assert t[0] in SETS["trace_width"] and t[1] in SETS["trace_lds_entries"]
        assert t[2] in SETS["trace_blocks_per_cu"] and t[3] in SETS["trace_block"] and t[4] in SETS["shade_tile"]
        assert t[5] > 0
    # the winner is the fastest trial, and it is the shape left set
    k = int(np.argmin([t[5] for t in trials]))
    assert res.best_ms == trials[k][5]
    assert trials[k][:5] == (res.best_width, res.best_lds, res.best_bpc, res.best_block, res.best_tile)
    # every axis was swept with the earlier axes at their chosen values
    assert all(t[0] == res.best_width for t in trials[2:])
    assert all(t[3] == res.best_block for t in trials[14:])
    # the chosen shape renders the default shape's frame
    seg, got = render(scene, img)
    for kk in SETS:
        ctx.tune(kk, {"trace_width": 4, "trace_lds_entries": 8, "trace_blocks_per_cu": 8, "trace_block": 256,
                      "shade_tile": 1024}[kk])
    seg0, ref = render(scene, img)
    assert seg == seg0 and np.allclose(got, ref, rtol=1e-5, atol=1e-6)
    scene.close()
