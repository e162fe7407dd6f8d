"""Writes tests/golden/ref_vectors.npz from the reference's OWN host-compiled
RayLib functions (oracle/_ref/libmrhref.so, built by `make -C oracle ref`
from the reference checkout).

Run in a checkout that has the reference:   python tests/golden/gen_golden.py

The fixture holds inputs and the reference's outputs only.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    O.build(reference=True)
    R = O.ref()
    rng = np.random.default_rng(20261202)
    out = {}

    # Morton: random triples + edges (0, max, and 2^bits which the mask wraps)
    q10 = rng.integers(0, 1024, size=(1000, 3)).astype(np.uint32)
    q10 = np.vstack([q10, [[0, 0, 0], [1023, 1023, 1023], [1024, 0, 0], [0, 1024, 0], [1, 2, 4]]]).astype(np.uint32)
    q21 = rng.integers(0, 1 << 21, size=(1000, 3)).astype(np.uint32)
    q21 = np.vstack([q21, [[0, 0, 0], [(1 << 21) - 1] * 3, [1 << 21, 0, 0], [5, 9, 17]]]).astype(np.uint32)
    out["morton30_q"] = q10
    out["morton30_code"] = np.array([R.ref_compose30(*map(int, v)) for v in q10], np.uint32)
    out["morton63_q"] = q21
    out["morton63_code"] = np.array([R.ref_compose63(*map(int, v)) for v in q21], np.uint64)

    # Moller-Trumbore: random rays/triangles, plus grazing / parallel / behind cases
    m = 1500
    o = rng.uniform(-1, 1, (m, 3)).astype(np.float32)
    d = rng.normal(size=(m, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tri = rng.uniform(-1, 1, (m, 9)).astype(np.float32)
    # aim half the rays at a point inside the triangle
    for i in range(0, m, 2):
        w = rng.dirichlet([1, 1, 1]).astype(np.float32)
        tgt = w[0] * tri[i, 0:3] + w[1] * tri[i, 3:6] + w[2] * tri[i, 6:9]
        v = tgt - o[i]
        d[i] = (v / np.linalg.norm(v)).astype(np.float32)
    # parallel ray in the triangle plane, ray starting on the triangle
    tri[1] = [0, 0, 0, 1, 0, 0, 0, 1, 0]; o[1] = [-1, 0.2, 0]; d[1] = [1, 0, 0]
    tri[3] = [0, 0, 0, 1, 0, 0, 0, 1, 0]; o[3] = [0.2, 0.2, 0]; d[3] = [0, 0, 1]
    hit = np.zeros(m, np.int32); t = np.zeros(m, np.float32); bary = np.zeros((m, 3), np.float32)
    for i in range(m):
        tt = np.zeros(1, np.float32); bb = np.zeros(3, np.float32)
        hit[i] = R.ref_intersect_triangle(p(o[i]), p(d[i]), p(tri[i]), p(tt), p(bb))
        t[i] = tt[0]; bary[i] = bb
    out.update(tri_o=o, tri_d=d, tri_v=tri, tri_hit=hit, tri_t=t, tri_bary=bary)

    # slab test
    bo = rng.uniform(-2, 2, (m, 3)).astype(np.float32)
    bd = rng.normal(size=(m, 3)); bd = (bd / np.linalg.norm(bd, axis=1, keepdims=True)).astype(np.float32)
    lo = rng.uniform(-1, 0.5, (m, 3)).astype(np.float32)
    box = np.hstack([lo, lo + rng.uniform(0.01, 1.0, (m, 3)).astype(np.float32)]).astype(np.float32)
    tr = np.stack([np.zeros(m), rng.uniform(0.5, 10, m)], 1).astype(np.float32)
    bh = np.zeros(m, np.int32); bt = np.zeros((m, 2), np.float32)
    for i in range(m):
        to = np.zeros(2, np.float32)
        bh[i] = R.ref_intersect_aabb(p(bo[i]), p(bd[i]), p(box[i]), float(tr[i, 0]), float(tr[i, 1]), p(to))
        bt[i] = to
    out.update(box_o=bo, box_d=bd, box_v=box, box_range=tr, box_hit=bh, box_t=bt)

    # Nudge (RT Gems ch. 6) at several magnitudes incl. the |p| < 1/32 branch
    pn = (rng.uniform(-3, 3, (m, 3)) * rng.choice([1.0, 1e-2, 1e-5], (m, 3))).astype(np.float32)
    nn = rng.normal(size=(m, 3)); nn = (nn / np.linalg.norm(nn, axis=1, keepdims=True)).astype(np.float32)
    no = np.zeros((m, 3), np.float32)
    for i in range(m):
        R.ref_nudge(p(pn[i]), p(nn[i]), p(nn[i]), p(no[i]))
    out.update(nudge_p=pn, nudge_n=nn, nudge_out=no)

    # RotationBetweenZAxis + ApplyRotation, incl. the two singular branches
    nq = rng.normal(size=(m, 3)); nq = (nq / np.linalg.norm(nq, axis=1, keepdims=True)).astype(np.float32)
    nq[0] = [0, 0, 1]; nq[1] = [0, 0, -1]; nq[2] = [0, 1e-5, 1]; nq[2] /= np.linalg.norm(nq[2])
    vq = rng.normal(size=(m, 3)).astype(np.float32)
    qq = np.zeros((m, 4), np.float32); rq = np.zeros((m, 3), np.float32)
    for i in range(m):
        R.ref_rot_between_z(p(nq[i]), p(qq[i]))
        R.ref_quat_rotate(p(qq[i]), p(vq[i]), p(rq[i]))
    out.update(rotz_n=nq, rotz_q=qq, rotz_v=vq, rotz_out=rq)

    # Hemisphere sampling
    xi = rng.random((m, 2)).astype(np.float32)
    xi[0] = [0, 0]; xi[1] = [np.float32(1 - 2 ** -24), np.float32(1 - 2 ** -24)]
    hu = np.zeros((m, 3), np.float32); hp = np.zeros(m, np.float32)
    for i in range(m):
        pdf = np.zeros(1, np.float32)
        R.ref_hemi_uniform(float(xi[i, 0]), float(xi[i, 1]), p(hu[i]), p(pdf)); hp[i] = pdf[0]
    out.update(hemi_xi=xi, hemi_dir=hu, hemi_pdf=hp)

    # Reflect, triangle normal, triangle AABB, centroid
    rw = rng.normal(size=(m, 3)).astype(np.float32)
    rr = np.zeros((m, 3), np.float32)
    for i in range(m):
        R.ref_reflect(p(rw[i]), p(nq[i]), p(rr[i]))
    tn = np.zeros((m, 3), np.float32); ta = np.zeros((m, 6), np.float32); tc = np.zeros((m, 3), np.float32)
    for i in range(m):
        R.ref_triangle_normal(p(tri[i]), p(tn[i]))
        R.ref_triangle_aabb(p(tri[i]), p(ta[i]))
        R.ref_aabb_centroid(p(ta[i]), p(tc[i]))
    out.update(reflect_wi=rw, reflect_n=nq, reflect_out=rr, trinorm_out=tn, triaabb_out=ta,
               centroid_out=tc)

    # Concentric octahedral map and its integer wrap (CoordinateConversion.h)
    r2 = np.random.default_rng(20261203)
    ouv = r2.random((m, 2)).astype(np.float32)
    ouv[:8] = [[0, 0], [1, 1], [0.5, 0.5], [0, 1], [1, 0], [0.5, 0], [0.25, 0.75], [0.5, 1]]
    odir = np.zeros((m, 3), np.float32)
    for i in range(m):
        R.ref_octa_to_dir(float(ouv[i, 0]), float(ouv[i, 1]), p(odir[i]))
    dd = r2.normal(size=(m, 3)); dd = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    dd[:4] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, -1, 0]]
    duv = np.zeros((m, 2), np.float32)
    for i in range(m):
        R.ref_dir_to_octa(p(dd[i]), p(duv[i]))
    wq = []
    for X, Y in ((8, 8), (16, 8), (64, 64), (32, 16)):
        for x in range(-6, X + 6):
            for y in (-6, -1, 0, Y // 2, Y - 1, Y, Y + 5):
                o2 = np.zeros(2, np.int32)
                R.ref_octa_wrap_int(x, y, X, Y, p(o2))
                wq.append((x, y, X, Y, int(o2[0]), int(o2[1])))
    out.update(octa_uv=ouv, octa_dir=odir, octa_in_dir=dd, octa_out_uv=duv,
               octa_wrap=np.array(wq, np.int32))

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
