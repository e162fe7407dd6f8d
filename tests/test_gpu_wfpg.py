"""Per-bin radiance fields of the reference's wavefront path guiding on the GPU
(mrh_wfpg_sample_fields) against the oracle (om_wfpg_sample_fields): filtered
fields, sampled (u,v), directions, sample pdfs and evaluated pdfs, bit for
bit, at every field size the reference's guided tracer uses (64, 32, 16, 8)
and the non-square sizes of its distribution test, over ragged bins (empty
bins, all-empty and uniform fields, a bright texel on the map's edge), ray
ids out of order or out of range, three filter widths; fed by
mrh_svo_bin_rays's partition; and at full size (4096 bins of 64x64 texels,
one ray per 1080p pixel) through size-independent properties plus a sample of
bins against the oracle."""
import math

import numpy as np
import pytest

from meturay_amd import native as N
from test_gpu_guide import dev, host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def make_case(rng, X, Y, bins, n_rays):
    fields = rng.gamma(0.6, 2.0, (bins, Y, X)).astype(np.float32)
    fields[rng.random((bins, Y, X)) < 0.15] = 0.0
    if bins > 3:
        fields[1] = 0.0                                   # all empty
        fields[2] = 1.0                                   # uniform
        fields[3, 0, 0] = 500.0                           # bright corner texel
        fields[3, Y // 2, X - 1] = 300.0                  # bright edge texel
    sizes = rng.integers(0, 2 * n_rays // bins + 1, bins)
    if bins > 4:
        sizes[4] = 0                                      # a bin without rays
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    total = int(offsets[-1])
    ids = rng.permutation(max(n_rays, total))[:total].astype(np.uint32)
    ids[rng.random(total) < 0.02] = n_rays + 5            # outside the ray arrays
    xi = rng.random((n_rays, 2)).astype(np.float32)
    xi[:4] = [[0, 0], [1 - 2 ** -24, 1 - 2 ** -24], [0.5, 0.5], [0, 1 - 2 ** -24]]
    ev = rng.random((n_rays, 2)).astype(np.float32)
    ev[:3] = [[0, 0], [1, 1], [0.999999, 0.0]]
    return fields, offsets, ids, xi, ev


def run_gpu(ctx, fields, offsets, ids, xi, ev, alpha, n_rays, dirs=True, filtered=True):
    bins, Y, X = fields.shape
    d_f = dev(fields)
    d_off = dev(offsets)
    d_ids = dev(ids if len(ids) else np.zeros(1, np.uint32))
    d_xi = dev(xi)
    d_ev = dev(ev) if ev is not None else None
    nan = float("nan")
    d_uv = torch.full((max(n_rays, 1), 2), nan, device="cuda")
    d_pdf = torch.full((max(n_rays, 1),), nan, device="cuda")
    d_epdf = torch.full((max(n_rays, 1),), nan, device="cuda") if ev is not None else None
    d_dir = torch.full((max(n_rays, 1), 3), nan, device="cuda") if dirs else None
    d_filt = torch.full((bins, Y, X), nan, device="cuda") if filtered else None
    torch.cuda.synchronize()
    ctx.wfpg_sample_fields(d_f, X, Y, bins, d_off, d_ids, n_rays, d_xi, d_uv, d_pdf, alpha=alpha, d_eval_uv=d_ev,
                           d_out_eval_pdf=d_epdf, d_out_dir=d_dir, d_out_filtered=d_filt)
    ctx.sync()
    out = {"uv": d_uv.cpu().numpy()[:n_rays], "pdf": d_pdf.cpu().numpy()[:n_rays]}
    out["eval_pdf"] = d_epdf.cpu().numpy()[:n_rays] if d_epdf is not None else None
    out["dir"] = d_dir.cpu().numpy()[:n_rays] if dirs else None
    out["filtered"] = d_filt.cpu().numpy() if filtered else None
    return out


def assert_same(got, want, keys=("uv", "pdf", "eval_pdf", "dir", "filtered")):
    for k in keys:
        if got[k] is None:
            continue
        g, w = got[k], want[k]
        assert g.shape == w.shape, k
        assert np.array_equal(g, w, equal_nan=True), (k, np.argwhere(~((g == w) | (np.isnan(g) & np.isnan(w))))[:5])


@pytest.mark.parametrize("X,Y", [(64, 64), (32, 32), (16, 16), (8, 8), (64, 32), (16, 8)])
def test_matches_oracle(ctx, oracle, X, Y):
    rng = np.random.default_rng(X * 7 + Y)
    n_rays = 5000
    fields, offsets, ids, xi, ev = make_case(rng, X, Y, 24, n_rays)
    print(f"[wfpg] {X}x{Y} bins 24 rays {n_rays}", flush=True)
    got = run_gpu(ctx, fields, offsets, ids, xi, ev, 1.0, n_rays)
    want = oracle.wfpg_sample_fields(fields, offsets, ids, xi, alpha=1.0, eval_uv=ev, ray_count=n_rays)
    assert_same(got, want)
    assert np.isfinite(got["pdf"]).sum() > 0.9 * len(ids)


@pytest.mark.parametrize("alpha", [0.25, 4.0, 1e-3])
def test_filter_widths(ctx, oracle, alpha):
    rng = np.random.default_rng(int(alpha * 1000) + 3)
    fields, offsets, ids, xi, ev = make_case(rng, 32, 32, 9, 2000)
    got = run_gpu(ctx, fields, offsets, ids, xi, None, alpha, 2000, dirs=False)
    want = oracle.wfpg_sample_fields(fields, offsets, ids, xi, alpha=alpha, ray_count=2000)
    assert_same(got, want, keys=("uv", "pdf", "filtered"))
    if alpha == 1e-3:   # a narrow Gaussian is the identity: taps (0,..,1,..,0)
        raised = np.where(fields == 0, np.float32(1e-4), fields)
        assert np.array_equal(got["filtered"], raised)


def test_optional_outputs_absent(ctx, oracle):
    rng = np.random.default_rng(8)
    fields, offsets, ids, xi, _ = make_case(rng, 16, 16, 6, 800)
    got = run_gpu(ctx, fields, offsets, ids, xi, None, 1.0, 800, dirs=False, filtered=False)
    want = oracle.wfpg_sample_fields(fields, offsets, ids, xi, alpha=1.0, ray_count=800)
    assert_same(got, want, keys=("uv", "pdf"))


def test_fed_by_ray_binning(ctx, oracle):
    """mrh_svo_bin_rays's partition (offsets, sorted ray ids) drives the fields."""
    rng = np.random.default_rng(21)
    cube, depth = [-2.0, -2.0, -2.0, 4.0], 6
    pts = (rng.random((20000, 3)) * 3.6 - 1.8).astype(np.float32)
    tree = oracle.svo_codes(pts, cube, depth)
    n_rays = 30000
    codes = oracle.svo_codes((rng.random((n_rays, 3)) * 3.6 - 1.8).astype(np.float32), cube, depth)
    svo = N.Svo(ctx)
    d_tree = dev(tree)
    torch.cuda.synchronize()
    svo.build(d_tree.data_ptr(), len(tree), cube, depth)
    d_codes = dev(codes)
    d_bins = torch.empty(n_rays, dtype=torch.int32, device="cuda")
    d_order = torch.empty_like(d_bins)
    d_off = torch.empty(n_rays + 1, dtype=torch.int32, device="cuda")
    d_pb = torch.empty_like(d_bins)
    d_cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    svo.bin_rays(d_codes.data_ptr(), n_rays, 2, 64, d_bins.data_ptr(), d_order.data_ptr(), d_off.data_ptr(),
                 d_pb.data_ptr(), d_cnt.data_ptr())
    ctx.sync()
    k = int(host(d_cnt, np.uint32)[0])
    assert k > 10
    fields = rng.gamma(0.6, 2.0, (k, 16, 16)).astype(np.float32)
    xi = rng.random((n_rays, 2)).astype(np.float32)
    d_f, d_xi = dev(fields), dev(xi)
    d_uv = torch.full((n_rays, 2), float("nan"), device="cuda")
    d_pdf = torch.full((n_rays,), float("nan"), device="cuda")
    torch.cuda.synchronize()
    ctx.wfpg_sample_fields(d_f, 16, 16, k, d_off, d_order, n_rays, d_xi, d_uv, d_pdf)
    ctx.sync()
    offsets = host(d_off, np.uint32)[:k + 1]
    order = host(d_order, np.uint32)[:n_rays]
    want = oracle.wfpg_sample_fields(fields, offsets, order[:offsets[-1]], xi, alpha=1.0, ray_count=n_rays)
    assert np.array_equal(d_uv.cpu().numpy(), want["uv"], equal_nan=True)
    assert np.array_equal(d_pdf.cpu().numpy(), want["pdf"], equal_nan=True)
    assert offsets[-1] == n_rays and not np.isnan(want["pdf"]).any()


def test_full_size_properties(ctx, oracle):
    """4096 bins of 64x64 texels, one ray per 1080p pixel."""
    X = Y = 64
    bins, n_rays = 4096, 1920 * 1080
    rng = np.random.default_rng(99)
    print(f"[wfpg] full size: bins {bins} rays {n_rays}", flush=True)
    fields = rng.gamma(0.5, 1.0, (bins, Y, X)).astype(np.float32)
    fields[rng.random((bins, Y, X)) < 0.3] = 0.0
    bin_of = np.sort(rng.integers(0, bins, n_rays))
    offsets = np.searchsorted(bin_of, np.arange(bins + 1)).astype(np.uint32)
    ids = rng.permutation(n_rays).astype(np.uint32)
    xi = rng.random((n_rays, 2)).astype(np.float32)
    got = run_gpu(ctx, fields, offsets, ids, xi, None, 1.0, n_rays, dirs=True, filtered=True)
    uv, pdf, dirs = got["uv"], got["pdf"], got["dir"]
    assert np.isfinite(uv).all() and (uv >= 0).all() and (uv <= 1).all()
    assert np.isfinite(pdf).all() and (pdf >= 0).all() and (pdf > 0).mean() > 0.9999
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-5)
    # idempotence: the pdf evaluated at each sample is the sample's pdf
    again = run_gpu(ctx, fields, offsets, ids, xi, uv, 1.0, n_rays, dirs=False, filtered=False)
    assert np.array_equal(again["uv"], uv) and np.array_equal(again["eval_pdf"], pdf)
    # filter keeps each field's total (after the empty-texel raise)
    raised = np.where(fields == 0, np.float32(1e-4), fields).astype(np.float64)
    ratio = got["filtered"].astype(np.float64).sum((1, 2)) / raised.sum((1, 2))
    assert np.abs(ratio - 1).max() < 1e-4
    # a sample of bins against the oracle, bit for bit
    for b in rng.choice(bins, 12, replace=False):
        lo, hi = int(offsets[b]), int(offsets[b + 1])
        one = np.array([0, hi - lo], np.uint32)
        want = oracle.wfpg_sample_fields(fields[b:b + 1], one, ids[lo:hi], xi, alpha=1.0, ray_count=n_rays)
        assert np.array_equal(got["filtered"][b], want["filtered"][0])
        r = ids[lo:hi]
        assert np.array_equal(uv[r], want["uv"][r]) and np.array_equal(pdf[r], want["pdf"][r])
        assert np.array_equal(dirs[r], want["dir"][r])
    # statistical: the mean of 1/pdf over each bin's samples estimates 4 pi
    # wherever the field has no empty texels after the raise (it has none)
    assert abs(np.mean(1.0 / pdf[pdf > 0]) / (4 * math.pi) - 1) < 0.02


def test_rejects_bad_arguments(ctx):
    f = torch.ones((2, 8, 8), device="cuda")
    off = dev(np.array([0, 1, 2], np.uint32))
    ids = dev(np.array([0, 1], np.uint32))
    xi = torch.zeros((2, 2), device="cuda")
    uv = torch.zeros((2, 2), device="cuda")
    pdf = torch.zeros(2, device="cuda")
    torch.cuda.synchronize()
    for X, Y, alpha in ((12, 8, 1.0), (8, 128, 1.0), (8, 8, 0.0), (8, 8, float("inf")), (8, 8, float("nan"))):
        with pytest.raises(N.MrhError):
            ctx.wfpg_sample_fields(f, X, Y, 2, off, ids, 2, xi, uv, pdf, alpha=alpha)
    with pytest.raises(N.MrhError):
        ctx.wfpg_sample_fields(f, 8, 8, 2, off, ids, 2, xi, uv, pdf, d_eval_uv=xi)
    with pytest.raises(N.MrhError):
        ctx.wfpg_sample_fields(f, 8, 8, 2, off, ids, 2, None, uv, pdf)
    ctx.wfpg_sample_fields(None, 8, 8, 0, None, None, 0, None, None, None)   # no bins: nothing to do
    ctx.sync()


@pytest.mark.parametrize("X,Y", [(64, 64), (32, 32), (16, 16), (8, 8), (64, 32)])
def test_product_mode_matches_oracle(ctx, oracle, X, Y):
    """The 8x8 product sampler: (u,v), direction, pdf and the pdf at a second
    (u,v), bit for bit, with per-ray BxDF cells (some near zero, so the 1e-3
    floor and long rejection runs occur)."""
    rng = np.random.default_rng(X + 3 * Y)
    n_rays = 3000
    fields, offsets, ids, xi, ev = make_case(rng, X, Y, 12, n_rays)
    bx = rng.gamma(0.5, 1.0, (n_rays, 64)).astype(np.float32)
    bx[rng.random((n_rays, 64)) < 0.2] = 0.0
    d_f, d_off, d_ids, d_ev, d_bx = dev(fields), dev(offsets), dev(ids), dev(ev), dev(bx.reshape(-1))
    nan = float("nan")
    d_uv = torch.full((n_rays, 2), nan, device="cuda")
    d_pdf = torch.full((n_rays,), nan, device="cuda")
    d_epdf = torch.full((n_rays,), nan, device="cuda")
    d_dir = torch.full((n_rays, 3), nan, device="cuda")
    torch.cuda.synchronize()
    ctx.wfpg_sample_fields(d_f, X, Y, 12, d_off, d_ids, n_rays, None, d_uv, d_pdf, alpha=1.0, d_eval_uv=d_ev,
                           d_out_eval_pdf=d_epdf, d_out_dir=d_dir, d_bxdf=d_bx, seed=1234)
    ctx.sync()
    want = oracle.wfpg_sample_fields(fields, offsets, ids, xi, alpha=1.0, eval_uv=ev, ray_count=n_rays, bxdf=bx,
                                     seed=1234)
    for k, t in (("uv", d_uv), ("pdf", d_pdf), ("eval_pdf", d_epdf), ("dir", d_dir)):
        assert np.array_equal(t.cpu().numpy(), want[k], equal_nan=True), k
    assert np.isfinite(want["pdf"]).sum() > 0.9 * len(ids)
