"""Statistical comparison of a render with a converged fixture.

Both sides are means over independent batches with a per-pixel standard
error (SE) of that mean, so the difference of two unbiased estimates of the
same image has variance se_a^2 + se_b^2 per pixel.  Pixels are independent
(each owns its random streams), so a t x t tile's mean has variance
sum(se^2) / t^4.

An absolute floor, 1e-3 of the reference image's mean radiance, is added in
quadrature to every pixel's and tile's standard error.  A plain path tracer
that found no light at a pixel in 4096 samples reports a mean and a standard
error of exactly zero there, although radiance too rare to be sampled is not
zero (next-event estimation finds 1e-5..1e-3 at such pixels); the floor keeps
those near-black pixels from deciding the test.

The checks and their tolerances (n = tiles):
  * deterministic pixels (both SEs 0: background, directly seen emitters):
    means equal within 1e-5 relative;
  * pixel z = diff / sqrt(se_a^2 + se_b^2 + floor^2): at most 0.2 % with |z| > 5
    (16-batch SEs make z roughly t-distributed with ~15 degrees of freedom,
    P(|t15| > 5) = 1.6e-4);
  * tile z: |mean| <= 5 / sqrt(n) (five standard errors of the mean of n
    unit normals), standard deviation in [0.75, 1.33], max |z| <= 6;
  * relMSE (eps 1e-2) at most twice its expectation mean(var / (ref^2 + eps)).
"""
from __future__ import annotations

import numpy as np


def batch_mean_se(batches):
    b = np.asarray(batches, np.float64)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(b.shape[0])


FLOOR = 1e-3


def compare(mean_a, se_a, mean_b, se_b, tile=8, eps=1e-2):
    mean_a = np.asarray(mean_a, np.float64)
    mean_b = np.asarray(mean_b, np.float64)
    var = np.asarray(se_a, np.float64) ** 2 + np.asarray(se_b, np.float64) ** 2
    floor2 = (FLOOR * float(np.mean(mean_b))) ** 2
    det = var == 0
    d = mean_a - mean_b
    out = {}
    out["det_pixels"] = int(det.sum())
    out["det_max_rel"] = float(np.max(np.abs(d[det]) / np.maximum(1.0, np.abs(mean_b[det])))) if det.any() else 0.0
    z = d[~det] / np.sqrt(var[~det] + floor2)
    out["pixels"] = int(z.size)
    out["pixel_frac_over5"] = float(np.mean(np.abs(z) > 5)) if z.size else 0.0
    h, w = mean_a.shape[:2]
    th, tw = h // tile, w // tile
    crop = (slice(0, th * tile), slice(0, tw * tile))

    def tiles(x):
        x = x[crop]
        return x.reshape(th, tile, tw, tile, -1).sum(axis=(1, 3))

    td = tiles(d) / tile ** 2
    tv = tiles(var) / tile ** 4
    ok = tv > 0
    zt = td[ok] / np.sqrt(tv[ok] + floor2)
    out["tiles"] = int(zt.size)
    out["tile_z_mean"] = float(zt.mean())
    out["tile_z_std"] = float(zt.std())
    out["tile_z_max"] = float(np.abs(zt).max())
    out["relmse"] = float(np.mean(d ** 2 / (mean_b ** 2 + eps)))
    out["relmse_expected"] = float(np.mean(var / (mean_b ** 2 + eps)))
    return out


def failures(r) -> list[str]:
    bad = []
    if r["det_max_rel"] > 1e-5:
        bad.append(f"deterministic pixels differ by {r['det_max_rel']:.3g}")
    if r["pixel_frac_over5"] > 0.002:
        bad.append(f"{r['pixel_frac_over5']:.4f} of pixels beyond 5 sigma")
    if abs(r["tile_z_mean"]) > 5.0 / np.sqrt(r["tiles"]):
        bad.append(f"tile z mean {r['tile_z_mean']:.3f} (bound {5.0 / np.sqrt(r['tiles']):.3f})")
    if not 0.75 <= r["tile_z_std"] <= 1.33:
        bad.append(f"tile z std {r['tile_z_std']:.3f}")
    if r["tile_z_max"] > 6.0:
        bad.append(f"tile z max {r['tile_z_max']:.2f}")
    if r["relmse"] > 2.0 * r["relmse_expected"]:
        bad.append(f"relMSE {r['relmse']:.3g} > 2 x expected {r['relmse_expected']:.3g}")
    return bad
