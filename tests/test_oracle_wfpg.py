"""Pins the oracle's per-bin radiance-field stage of the reference's wavefront
path guiding (oracle/om_wfpg.c; WFPGTracerKC.cuh:1252-1545).

- the concentric octahedral map and its integer edge fold against the
  reference's own host-compiled CoordinateConversion.h (golden vectors);
- the fold against the map's geometry: a fold through an edge lands on the
  texel that shares that edge point on the sphere;
- the Gaussian taps against their closed form;
- the filter against a float64 brute force and closed forms (constant field,
  interior impulse);
- the piecewise-constant distribution against the reference's own test
  (Test_TracerCUDA/T_BlockPWCDistribution.cu: all-ones known answer, random
  fields within 1e-4 of a float64 restatement at its six sizes);
- sampling against an inverse-cdf brute force (numpy searchsorted), the
  uniform field, zero rows, the ends of [0,1), and a chi-square test;
- the whole stage against the composition of its parts on ragged bins."""
import numpy as np
import pytest
from scipy import stats

SIZES = [(64, 64), (64, 32), (32, 32), (32, 16), (16, 16), (16, 8), (8, 8)]   # (X, Y)


@pytest.fixture(scope="module")
def golden():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


# ---------------------------------------------------------------- octahedral map

def test_octa_map_matches_reference(oracle, golden):
    d = oracle.octa_to_dirs(golden["octa_uv"])
    assert np.abs(d - golden["octa_dir"]).max() < 2e-6
    src = golden["octa_in_dir"]
    uv = oracle.dir_to_octas(src)
    pole = (src[:, 0] == 0) & (src[:, 1] == 0)
    assert np.abs(uv[~pole] - golden["octa_out_uv"][~pole]).max() < 2e-6
    # the reference sends both poles to (0,0), a corner that decodes to -z; the
    # oracle's +z pole sits at the centre.  Only -z is comparable, and it is
    # the same point of the sphere.
    for i in np.where(pole)[0]:
        a = oracle.octa_to_dirs(uv[i:i + 1])[0]
        assert np.allclose(a, src[i], atol=1e-6)


def test_octa_fold_matches_reference(oracle, golden):
    rows = golden["octa_wrap"]
    assert len(rows) > 1000
    for x, y, X, Y, ex, ey in rows:
        assert oracle.octa_wrap_int(int(x), int(y), int(X), int(Y)) == (int(ex), int(ey)), (x, y, X, Y)


@pytest.mark.parametrize("X,Y", [(8, 8), (16, 8), (32, 32)])
def test_octa_fold_follows_the_map_seams(oracle, X, Y):
    """A fold through an edge lands on the texel across that edge on the sphere:
    the point where texel (0, y) meets its u-neighbour must be the point where
    the folded texel meets the map edge, and likewise for v."""
    for y in range(Y):
        fx, fy = oracle.octa_wrap_int(-1, y, X, Y)
        assert fx == 0
        a = oracle.octa_to_dirs(np.array([[0.0, (y + 0.5) / Y]], np.float32))[0]
        b = oracle.octa_to_dirs(np.array([[0.0, (fy + 0.5) / Y]], np.float32))[0]
        assert np.allclose(a, b, atol=2e-6), (y, fy)
        fx, fy = oracle.octa_wrap_int(X, y, X, Y)
        a = oracle.octa_to_dirs(np.array([[1.0, (y + 0.5) / Y]], np.float32))[0]
        b = oracle.octa_to_dirs(np.array([[1.0, (fy + 0.5) / Y]], np.float32))[0]
        assert fx == X - 1 and np.allclose(a, b, atol=2e-6)
    for x in range(X):
        fx, fy = oracle.octa_wrap_int(x, -1, X, Y)
        a = oracle.octa_to_dirs(np.array([[(x + 0.5) / X, 0.0]], np.float32))[0]
        b = oracle.octa_to_dirs(np.array([[(fx + 0.5) / X, 0.0]], np.float32))[0]
        assert fy == 0 and np.allclose(a, b, atol=2e-6)


# ---------------------------------------------------------------- Gaussian filter

@pytest.mark.parametrize("alpha", [0.25, 1.0, 4.0])
def test_gauss_taps_closed_form(oracle, alpha):
    w = oracle.wfpg_gauss_weights(alpha)
    k = np.arange(-5, 6, dtype=np.float64)
    g = np.exp(-k * k / (2 * alpha))
    assert np.allclose(w, g / g.sum(), rtol=2e-6, atol=1e-9)
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    assert np.array_equal(w, w[::-1])


def _filter64(oracle, f, alpha):
    Y, X = f.shape
    w = oracle.wfpg_gauss_weights(alpha).astype(np.float64)
    a = f.astype(np.float64)
    t = np.zeros_like(a)
    for y in range(Y):
        for x in range(X):
            s = 0.0
            for j in range(-5, 6):
                qx, qy = x + j, y
                if qx < 0 or qx >= X:
                    qx, qy = oracle.octa_wrap_int(qx, qy, X, Y)
                s += a[qy, qx] * w[j + 5]
            t[y, x] = s
    o = np.zeros_like(a)
    for y in range(Y):
        for x in range(X):
            s = 0.0
            for j in range(-5, 6):
                qx, qy = x, y + j
                if qy < 0 or qy >= Y:
                    qx, qy = oracle.octa_wrap_int(qx, qy, X, Y)
                s += t[qy, qx] * w[j + 5]
            o[y, x] = s
    return o


@pytest.mark.parametrize("X,Y", [(8, 8), (16, 8), (32, 16)])
def test_filter_matches_float64_brute_force(oracle, X, Y):
    rng = np.random.default_rng(X * 100 + Y)
    f = rng.uniform(0, 10, (Y, X)).astype(np.float32)
    f[rng.random((Y, X)) < 0.2] *= 50          # a few bright texels near edges too
    got = oracle.wfpg_filter(f, 1.0)
    ref = _filter64(oracle, f, 1.0)
    assert np.allclose(got, ref, rtol=2e-6, atol=0)
    # the fold is an involution of the border and the taps are symmetric:
    # the filter keeps the field's total
    assert abs(got.astype(np.float64).sum() / f.astype(np.float64).sum() - 1) < 1e-5


@pytest.mark.parametrize("X,Y", SIZES)
def test_filter_closed_forms(oracle, X, Y):
    c = np.full((Y, X), 3.5, np.float32)
    assert np.allclose(oracle.wfpg_filter(c, 1.0), 3.5, rtol=2e-6)
    if X >= 16 and Y >= 16:
        imp = np.zeros((Y, X), np.float32)
        imp[Y // 2, X // 2] = 1.0
        got = oracle.wfpg_filter(imp, 2.0)
        w = oracle.wfpg_gauss_weights(2.0)
        want = np.zeros((Y, X), np.float32)
        want[Y // 2 - 5:Y // 2 + 6, X // 2 - 5:X // 2 + 6] = np.outer(w, w)
        assert np.array_equal(got, want)


# ---------------------------------------------------------------- distribution

@pytest.mark.parametrize("X,Y", SIZES)
def test_pwc_all_ones_known_answer(oracle, X, Y):
    """T_BlockPWCDistribution.cu BasicInit: pdf 1 everywhere, cdf i / X."""
    d = oracle.Pwc2D(np.ones((Y, X), np.float32))
    ulp = 4 * np.finfo(np.float32).eps
    assert np.allclose(d.pdfx, 1.0, rtol=ulp, atol=0)
    assert np.allclose(d.pdfy, 1.0, rtol=ulp, atol=0)
    assert np.allclose(d.cdfx, np.tile(np.arange(1, X + 1) / X, (Y, 1)), rtol=ulp, atol=0)
    assert np.allclose(d.cdfy, np.arange(1, Y + 1) / Y, rtol=ulp, atol=0)


@pytest.mark.parametrize("X,Y", SIZES)
def test_pwc_random_fields_within_reference_tolerance(oracle, X, Y):
    """T_BlockPWCDistribution.cu Stress: uniform [0,10) fields, every pdf/cdf
    within LargeEpsilon (1e-4) of a float64 restatement."""
    rng = np.random.default_rng(7 + X + Y)
    for _ in range(10):
        f = rng.uniform(0, 10, (Y, X)).astype(np.float32)
        d = oracle.Pwc2D(f)
        f64 = f.astype(np.float64)
        rows = f64.sum(1)
        assert np.allclose(d.pdfx, f64 / rows[:, None] * X, atol=1e-4, rtol=0)
        assert np.allclose(d.cdfx, np.cumsum(f64, 1) / rows[:, None], atol=1e-4, rtol=0)
        assert np.allclose(d.pdfy, rows / rows.sum() * Y, atol=1e-4, rtol=0)
        assert np.allclose(d.cdfy, np.cumsum(rows) / rows.sum(), atol=1e-4, rtol=0)
        assert abs((d.pdfx * d.pdfy[:, None]).astype(np.float64).mean() - 1) < 1e-5


def test_sample_uniform_field(oracle):
    d = oracle.Pwc2D(np.ones((32, 64), np.float32))
    rng = np.random.default_rng(3)
    for xi in rng.random((500, 2)).astype(np.float32):
        uv = d.sample(float(xi[0]), float(xi[1]))
        assert np.allclose(uv, xi, atol=1e-6)
        assert abs(d.pdf(float(uv[0]), float(uv[1])) * 4 * np.pi - 1) < 1e-6


@pytest.mark.parametrize("X,Y", [(8, 8), (32, 16), (64, 64)])
def test_sample_is_inverse_cdf(oracle, X, Y):
    rng = np.random.default_rng(11 * X + Y)
    f = rng.uniform(0.1, 10, (Y, X)).astype(np.float32)
    f[rng.random((Y, X)) < 0.1] = 0.0
    f[1] = 0.0                                    # an empty row
    d = oracle.Pwc2D(f)
    f64 = f.astype(np.float64)
    cy = np.cumsum(f64.sum(1)) / f64.sum()
    checked = 0
    for xi0, xi1 in rng.random((3000, 2)).astype(np.float32):
        uv = d.sample(float(xi0), float(xi1))
        iy = int(np.searchsorted(cy, xi1, side="right"))
        cx = np.cumsum(f64[iy]) / f64[iy].sum()
        ix = int(np.searchsorted(cx, xi0, side="right"))
        if np.min(np.abs(cy - xi1)) < 1e-6 or np.min(np.abs(cx - xi0)) < 1e-6:
            continue
        checked += 1
        assert int(uv[1] * Y) == iy and int(uv[0] * X) == ix
        lo = cx[ix - 1] if ix else 0.0
        # float32 cdf entries carry ~1e-7 of rounding; a narrow texel magnifies it
        assert abs(uv[0] * X - ix - (xi0 - lo) / (cx[ix] - lo)) < 1e-5 + 4e-7 / (cx[ix] - lo)
        p = d.pdf(float(uv[0]), float(uv[1]))
        assert p > 0 and abs(p - f64[iy, ix] / f64.sum() * X * Y / (4 * np.pi)) < 1e-4 * p + 1e-7
    assert checked > 2900
    # the ends of [0,1): the first and last texels with mass
    nz = np.argwhere(f > 0)
    uv = d.sample(0.0, 0.0)
    first_row = nz[:, 0].min()
    assert int(uv[1] * Y) == first_row and int(uv[0] * X) == np.argwhere(f[first_row] > 0).min()
    # at the top a float cdf that ends at or below xi clamps to the last texel,
    # empty or not ("exceeds the cdf: clamp to edge", GPUBlockPWCDistribution.cuh:240-246)
    top = np.float32(1 - 2 ** -24)
    uv = d.sample(float(top), float(top))
    # (the continuous index k + frac + 1 may round up onto the texel's end)
    row = Y - 1 if top >= d.cdfy[-1] else nz[:, 0].max()
    col = X - 1 if top >= d.cdfx[row, -1] else np.argwhere(f[row] > 0).max()
    assert int(uv[1] * Y) == row and (int(uv[0] * X) == col or uv[0] * X == col + 1)


def test_sample_chi_square(oracle):
    rng = np.random.default_rng(5)
    f = rng.uniform(0.05, 4, (8, 8)).astype(np.float32)
    d = oracle.Pwc2D(f)
    n = 60000
    counts = np.zeros((8, 8))
    for xi0, xi1 in rng.random((n, 2)).astype(np.float32):
        uv = d.sample(float(xi0), float(xi1))
        counts[int(uv[1] * 8), int(uv[0] * 8)] += 1
    p = f.astype(np.float64) / f.astype(np.float64).sum()
    assert stats.chisquare(counts.ravel(), p.ravel() * n).pvalue > 1e-4


# ---------------------------------------------------------------- the whole stage

def test_stage_is_the_composition_of_its_parts(oracle):
    rng = np.random.default_rng(19)
    X, Y, bins, n = 16, 16, 4, 300
    fields = rng.uniform(0, 5, (bins, Y, X)).astype(np.float32)
    fields[1, 3:6] = 0.0                          # empty texels take 1e-4
    fields[2] = 0.0                               # an all-empty field
    offsets = np.array([0, 90, 90, 200, 260], np.uint32)   # bin 1 has no rays
    ids = rng.permutation(n).astype(np.uint32)[:260]
    ids[5] = n + 7                                # outside ray_count: skipped
    xi = rng.random((n, 2)).astype(np.float32)
    ev = rng.random((n, 2)).astype(np.float32)
    out = oracle.wfpg_sample_fields(fields, offsets, ids, xi, alpha=1.0, eval_uv=ev, ray_count=n)
    seen = np.zeros(n, bool)
    for b in range(bins):
        raw = np.where(fields[b] == 0, np.float32(1e-4), fields[b]).astype(np.float32)
        filt = oracle.wfpg_filter(raw, 1.0)
        assert np.array_equal(out["filtered"][b], filt)
        d = oracle.Pwc2D(filt)
        for r in ids[offsets[b]:offsets[b + 1]]:
            if r >= n:
                continue
            seen[r] = True
            uv = d.sample(float(xi[r, 0]), float(xi[r, 1]))
            assert np.array_equal(out["uv"][r], uv)
            assert out["pdf"][r] == np.float32(d.pdf(float(uv[0]), float(uv[1])))
            assert out["eval_pdf"][r] == np.float32(d.pdf(float(ev[r, 0]), float(ev[r, 1])))
            z = oracle.octa_to_dirs(uv[None])[0]
            assert np.array_equal(out["dir"][r], np.array([z[1], z[2], z[0]], np.float32))
    assert np.isnan(out["uv"][~seen]).all() and np.isnan(out["pdf"][~seen]).all()
    # an all-empty field is uniform after the raise
    r = ids[offsets[2]]
    assert abs(out["pdf"][r] * 4 * np.pi - 1) < 1e-5


# ---------------------------------------------------------------- product mode

def _one_bin(field, n, bxdf_row, seed=3, ev=None):
    import numpy as _np
    bins = field[None]
    off = _np.array([0, n], _np.uint32)
    ids = _np.arange(n, dtype=_np.uint32)
    bx = _np.tile(bxdf_row.astype(_np.float32), (n, 1))
    return bins, off, ids, bx


def test_product_uniform_field_and_bxdf(oracle):
    f = np.ones((16, 16), np.float32)
    bins, off, ids, bx = _one_bin(f, 400, np.ones(64))
    out = oracle.wfpg_sample_fields(bins, off, ids, np.zeros((400, 2), np.float32), alpha=1e-3, bxdf=bx, seed=5)
    assert np.allclose(out["pdf"] * 4 * np.pi, 1.0, atol=1e-5)
    assert ((out["uv"] >= 0) & (out["uv"] < 1)).all()


def test_product_samples_follow_bxdf_times_field(oracle):
    """Texel frequencies match bxdf[cell] x field[texel] (chi-square), and each
    sample's pdf equals the pdf evaluated at it."""
    rng = np.random.default_rng(12)
    X = Y = 16
    f = rng.uniform(0.2, 3.0, (Y, X)).astype(np.float32)
    bxdf = rng.uniform(0.5, 2.0, 64).astype(np.float32)
    n = 60000
    bins, off, ids, bx = _one_bin(f, n, bxdf)
    # alpha 1e-3 makes the filter the identity, so the sampled field is f
    out = oracle.wfpg_sample_fields(bins, off, ids, np.zeros((n, 2), np.float32), alpha=1e-3, bxdf=bx, seed=9)
    uv = out["uv"]
    ev = oracle.wfpg_sample_fields(bins, off, ids, np.zeros((n, 2), np.float32), alpha=1e-3, bxdf=bx, seed=9,
                                   eval_uv=uv)
    assert np.allclose(ev["eval_pdf"], out["pdf"], rtol=1e-6)
    tx = np.minimum((uv[:, 0] * X).astype(int), X - 1)
    ty = np.minimum((uv[:, 1] * Y).astype(int), Y - 1)
    counts = np.zeros((Y, X))
    np.add.at(counts, (ty, tx), 1)
    cell = (np.arange(Y)[:, None] // 2) * 8 + (np.arange(X)[None, :] // 2)
    want = bxdf[cell].astype(np.float64) * f.astype(np.float64)
    want /= want.sum()
    assert stats.chisquare(counts.ravel(), want.ravel() * n).pvalue > 1e-4
    # the pdf integrates to one over the sphere
    cu = ((np.arange(X) + 0.5) / X).astype(np.float32)
    grid = np.stack(np.meshgrid(cu, cu), -1).reshape(-1, 2).astype(np.float32)
    m = len(grid)
    b2, o2, i2, x2 = _one_bin(f, m, bxdf)
    g = oracle.wfpg_sample_fields(b2, o2, i2, np.zeros((m, 2), np.float32), alpha=1e-3, bxdf=x2, seed=9,
                                  eval_uv=grid)
    assert abs(float(g["eval_pdf"].astype(np.float64).sum()) * 4 * np.pi / m - 1.0) < 1e-4
