"""Writes the converged-image fixtures: 4096 spp per pixel rendered by the
oracle (the C restatement of the path tracer, radiance accumulated in FP64),
as 16 independent batches of 256 spp, so each pixel carries its mean and
the standard error of that mean.

    python tests/golden/gen_converged.py          (about 3 minutes on 8 cores)

Scenes:
  c1     the Cornell box of the 256x256 plain path-tracing workload: 12-triangle
         box, 5120-triangle icosphere, 2-triangle light of radiance 15; depth 8,
         Russian roulette after depth 3.
  grid8  an 8x8 grid of displaced UV spheres (16x16 tessellation, seed 2) with
         Lambert, GGX conductor, dielectric and emissive spheres; 96x64,
         depth 12, Russian roulette after depth 3 (the guided workload's
         material mix at a size the oracle converges in seconds).

The fixture holds per-pixel means and standard errors only (float32).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from meturay_amd import native as N  # noqa: E402  (host-only scene generators)
from oracle import oracle as O  # noqa: E402

BATCHES = 16
STRATA = 16             # 256 spp per batch, 4096 in total
SEED0 = 0x5EED0000      # oracle batches use SEED0 + b; tests use other seeds

SCENES = {
    "c1": dict(gen=lambda: N.gen_cornell(4), w=256, h=256, depth=8, rr=3),
    "grid8": dict(gen=lambda: N.gen_sphere_grid(8, 16, 16, 2), w=96, h=64, depth=12, rr=3),
}


def render_batches(name, threads):
    cfg = SCENES[name]
    sc = cfg["gen"]()
    w, h = cfg["w"], cfg["h"]
    bvh = O.LBVH(sc["tris"], 30)
    c = sc["camera"]
    cam = O.camera(c["position"], c["gaze"], c["up"], c["fov_x"], c["near"], c["far"], w, h)
    imgs = []
    for b in range(BATCHES):
        img, _ = O.render(sc, bvh, cam, w, h, STRATA, cfg["depth"], cfg["rr"], SEED0 + b,
                          threads=threads)
        imgs.append(img)
    imgs = np.stack(imgs)                      # [B, h, w, 3] float64
    mean = imgs.mean(axis=0)
    se = imgs.std(axis=0, ddof=1) / np.sqrt(BATCHES)
    return mean, se


def main():
    threads = os.cpu_count() or 8
    for name in SCENES:
        t0 = time.time()
        mean, se = render_batches(name, threads)
        path = os.path.join(HERE, f"converged_{name}_4096spp.npz")
        np.savez_compressed(path, mean=mean.astype(np.float32), se=se.astype(np.float32),
                            spp=np.int32(BATCHES * STRATA * STRATA), batches=np.int32(BATCHES),
                            seed0=np.uint32(SEED0))
        print(f"{name}: {time.time() - t0:.1f} s -> {path}", flush=True)


if __name__ == "__main__":
    main()
