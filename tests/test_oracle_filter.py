"""Reconstruction filters in the oracle, pinned by closed forms, a partition
of unity, and a brute-force scatter restatement of the reference's
KCExpandSamplesToPixels + segmented reduction.

- FilterRadiusToPixelWH follows the table in its own comment
  (GPUReconFilterKC.cuh:11-19);
- Box / Tent / Gaussian weights equal their closed forms; the Mitchell
  kernel takes the published values k(0) = (6 - 2B) / 6, k(1) = B / 6 and
  its integer translates sum to one when B + 2C = 1 (so does the tent at
  radius 1);
- the gather equals a per-sample scatter over every radius / type;
- filtering constant samples gives the constant; Box radius 0 is the
  per-pixel mean.
"""
import math

import numpy as np
import pytest

from meturay_amd import native as N

TYPES = ["Box", "Tent", "Gaussian", "Mitchell"]


def desc(oracle, kind, r):
    if kind == "Gaussian":
        return oracle.filter_desc(kind, r, alpha=2.0)
    if kind == "Mitchell":
        return oracle.filter_desc(kind, r, b=1.0 / 3.0, c=1.0 / 3.0)
    return oracle.filter_desc(kind, r)


def test_pixel_window_table(oracle):
    table = {0.0: 1, 0.1: 2, 0.5: 2, 0.5001: 3, 1.0: 3, 1.2: 4, 1.5: 4, 1.75: 5, 2.0: 5, 3.0: 7}
    for r, wh in table.items():
        assert oracle.filter_pixel_wh(r) == wh, r


def test_weights_closed_forms(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        r = float(rng.uniform(0.3, 2.5))
        dx, dy = (float(v) for v in rng.uniform(0, 1.2 * r, 2))
        box = oracle.filter_weight(oracle.filter_desc("Box", r), dx, dy)
        assert box == (1.0 if (dx < r and dy < r) else 0.0)
        tent = oracle.filter_weight(oracle.filter_desc("Tent", r), dx, dy)
        assert tent == pytest.approx(max(0, 1 - dx / r) * max(0, 1 - dy / r), rel=1e-5, abs=1e-7)
        a = 1.7
        g = oracle.filter_weight(oracle.filter_desc("Gaussian", r, alpha=a), dx, dy)
        e = math.exp(-a * r * r)
        assert g == pytest.approx(max(0, math.exp(-a * dx * dx) - e) * max(0, math.exp(-a * dy * dy) - e),
                                  rel=1e-5, abs=1e-7)
    assert oracle.filter_weight(oracle.filter_desc("Box", 0.0), 3.0, 5.0) == 1.0


def mitchell_1d(oracle, x, b, c):
    """1D kernel k(x) through the 2D weight at radius 2 (argument 2|d|/r = |d|)
    divided by k(0)."""
    d = oracle.filter_desc("Mitchell", 2.0, b=b, c=c)
    return oracle.filter_weight(d, x, 0.0) / oracle.filter_weight(d, 0.0, 0.0)


@pytest.mark.parametrize("b,c", [(1 / 3, 1 / 3), (0.0, 0.5), (1.0, 0.0), (0.5, 0.25)])
def test_mitchell_published_kernel(oracle, b, c):
    d = oracle.filter_desc("Mitchell", 2.0, b=b, c=c)
    k0 = oracle.filter_weight(d, 0.0, 0.0) ** 0.5      # k(0)^2 at the centre
    assert k0 == pytest.approx((6 - 2 * b) / 6, rel=1e-6)
    k1 = oracle.filter_weight(d, 1.0, 0.0) / k0
    assert k1 == pytest.approx(b / 6, rel=1e-5, abs=1e-7)
    # continuous at |x| = 1 and zero from |x| = 2 on
    lo = oracle.filter_weight(d, 1.0 - 1e-4, 0.0) / k0
    hi = oracle.filter_weight(d, 1.0 + 1e-4, 0.0) / k0
    assert abs(lo - hi) < 1e-3
    assert oracle.filter_weight(d, 2.0, 0.0) == 0.0
    # integer translates sum to one when B + 2C = 1
    if abs(b + 2 * c - 1) < 1e-9:
        for x in np.linspace(0.0, 1.0, 11):
            tot = sum(oracle.filter_weight(d, abs(x + k), 0.0) / k0 for k in range(-3, 4))
            assert tot == pytest.approx(1.0, abs=1e-5)


def test_tent_partition_of_unity(oracle):
    d = oracle.filter_desc("Tent", 1.0)
    for x in np.linspace(0.0, 1.0, 11):
        tot = sum(oracle.filter_weight(d, abs(x + k), 0.0) for k in range(-2, 3))
        assert tot == pytest.approx(1.0, abs=1e-6)


def scatter(oracle, d, values, coords, w, h):
    """The reference's algorithm restated per sample (GPUReconFilterKC.cuh:29-128
    expansion, then a per-pixel reduction), in float64."""
    r = d.radius
    rng_ = oracle.filter_pixel_wh(r)
    lo, hi = -((rng_ - 1) // 2), (rng_ + 2) // 2
    wsum = np.zeros((h, w))
    vsum = np.zeros((h, w, 3))
    for s in range(len(values)):
        cx, cy = float(coords[s, 0]), float(coords[s, 1])
        if not (0 <= cx < w and 0 <= cy < h):
            continue
        ix, iy = math.floor(cx), math.floor(cy)
        rx, ry = cx - ix, cy - iy
        xr, yr = [lo, hi], [lo, hi]
        if rng_ % 2 == 0:
            if rx < 0.5:
                xr = [lo - 1, hi - 1]
            if ry < 0.5:
                yr = [lo - 1, hi - 1]
        for y in range(*yr):
            for x in range(*xr):
                pcx, pcy = ix + x + 0.5, iy + y + 0.5
                if rng_ != 1 and (cx - pcx) ** 2 + (cy - pcy) ** 2 > r * r:
                    continue
                px, py = ix + x, iy + y
                if not (0 <= px < w and 0 <= py < h):
                    continue
                wt = oracle.filter_weight(d, abs(pcx - cx), abs(pcy - cy))
                wsum[py, px] += wt
                vsum[py, px] += wt * values[s]
    return vsum, wsum


@pytest.mark.parametrize("kind", TYPES)
@pytest.mark.parametrize("r", [0.0, 0.5, 0.75, 1.0, 1.5, 2.0])
def test_gather_equals_scatter(oracle, kind, r):
    if r == 0.0 and kind != "Box":
        pytest.skip("radius 0 is the Box filter's")
    rng = np.random.default_rng(int(r * 10) + TYPES.index(kind))
    w, h, n = 12, 9, 700
    coords = np.stack([rng.uniform(-0.5, w + 0.5, n), rng.uniform(-0.5, h + 0.5, n)], 1).astype(np.float32)
    coords[:20] = np.floor(coords[:20]) + 0.5        # centred samples: ties on the radius
    # samples on pixel edges: at half-integer radii the pixel two (or one)
    # below is exactly at the radius, inside the shifted window of even widths
    coords[20:40] = np.floor(coords[20:40]) + np.array([0.0, 0.5], np.float32)
    coords[40:60] = np.floor(coords[40:60]) + np.array([0.5, 0.0], np.float32)
    coords[60:70] = np.floor(coords[60:70])
    values = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    d = desc(oracle, kind, r)
    weighted, weights = oracle.filter_samples(d, values, coords, w, h)
    vs, ws = scatter(oracle, d, values, coords, w, h)
    assert np.allclose(weights, ws, rtol=1e-5, atol=1e-6)
    assert np.allclose(weighted, vs, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("kind", TYPES)
def test_constant_samples_resolve_to_constant(oracle, kind):
    rng = np.random.default_rng(9)
    w, h, spp = 10, 8, 9
    n = w * h * spp
    pix = np.repeat(np.arange(w * h), spp)
    coords = np.stack([pix % w + rng.random(n), pix // w + rng.random(n)], 1).astype(np.float32)
    values = np.full((n, 3), 0.7, np.float32)
    weighted, weights = oracle.filter_samples(desc(oracle, kind, 1.5), values, coords, w, h, grouped_spp=spp)
    img = oracle.filter_resolve(weighted, weights)
    ok = weights != 0
    assert ok.mean() > 0.9
    assert np.allclose(img[ok], 0.7, rtol=2e-6)
    # grouped order given explicitly sorts the same way: bit-identical
    weighted2, weights2 = oracle.filter_samples(desc(oracle, kind, 1.5), values, coords, w, h)
    assert np.array_equal(weighted, weighted2) and np.array_equal(weights, weights2)


def test_box_zero_is_the_pixel_mean(oracle):
    sc = N.gen_cornell(2)
    bvh = oracle.LBVH(sc["tris"], 30)
    c = sc["camera"]
    w, h, S = 24, 20, 3
    cam = oracle.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    plain, seg = oracle.render(sc, bvh, cam, w, h, S, 6, 3, 5, threads=8)
    filt, seg2 = oracle.render_filtered(sc, bvh, cam, w, h, S, 6, 3, 5, oracle.filter_desc("Box", 0.0), threads=8)
    assert seg == seg2
    assert np.allclose(filt, plain, rtol=1e-6, atol=1e-7)
    # a wider filter keeps the mean level and smooths
    tent, _ = oracle.render_filtered(sc, bvh, cam, w, h, S, 6, 3, 5, oracle.filter_desc("Tent", 1.0), threads=8)
    assert abs(tent.mean() / plain.mean() - 1) < 0.05
    assert np.abs(np.diff(tent, axis=1)).mean() < np.abs(np.diff(plain, axis=1)).mean()
