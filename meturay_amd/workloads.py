"""Request streams of the serving workload (host only, numpy).

The stream: Poisson arrivals, 8 resident sphere-grid scenes whose triangle
counts are log-uniform in [10K, 4M] (grids of g x g UV spheres with 16x16
tessellation, 512 triangles each, scene k generated with seed 100 + k),
resolution from {256x256, 512x512, 1024x768, 1920x1080}, spp from
{1, 4, 16, 64} and guiding on with probability 0.5 (a guided request renders
min(4, spp) learn-then-render iterations of spp / iterations samples each).
"""
from __future__ import annotations

import numpy as np

SERVE_RES = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
SERVE_SPP = [1, 4, 16, 64]
SCENE_STACKS = SCENE_SLICES = 16


def serving_workload(n_req: int, rate: float, seed: int = 3):
    """Returns (grid sizes of the 8 scenes, requests).  Each request is a
    dict: t (arrival, s), scene, w, h, S (strata per axis of an iteration),
    it (iterations; 1 = plain), spp, seed (= its index)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), 8))
    grids = [int(max(5, round(np.sqrt(x / 512.0)))) for x in sizes]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for i in range(n_req):
        w, h = SERVE_RES[rng.integers(0, 4)]
        spp = SERVE_SPP[rng.integers(0, 4)]
        guided = bool(rng.random() < 0.5)
        it = min(4, spp) if guided else 1
        S = int(round(np.sqrt(spp // it)))
        reqs.append({"t": float(arrivals[i]), "scene": int(rng.integers(0, 8)), "w": w, "h": h, "S": S,
                     "it": it, "spp": spp, "seed": i})
    return grids, reqs


def serving_scene(N, grid: int, k: int):
    """Scene k of the stream: a grid x grid sphere grid, seed 100 + k."""
    return N.gen_sphere_grid(grid, SCENE_STACKS, SCENE_SLICES, 100 + k)
