"""Pins the oracle against golden vectors written by the reference's OWN
host-compiled RayLib functions (tests/golden/gen_golden.py -> ref_vectors.npz).

Functions restated with the reference's operation order must match bit for
bit; the hemisphere sampler differs only in how sin/cos(2*pi*u) is evaluated
and must match to 1e-6.
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def test_morton_compose_matches_reference(oracle, g):
    L = oracle.lib()
    got30 = np.array([L.om_morton_compose30(*map(int, q)) for q in g["morton30_q"]], np.uint32)
    got63 = np.array([L.om_morton_compose63(*map(int, q)) for q in g["morton63_q"]], np.uint64)
    assert np.array_equal(got30, g["morton30_code"])
    assert np.array_equal(got63, g["morton63_code"])
    # the 2^bits inputs wrap to 0 through the reference's mask
    assert L.om_morton_compose30(1024, 0, 0) == 0


def test_triangle_intersection_bit_exact(oracle, g):
    L = oracle.lib()
    o, d, tri = g["tri_o"], g["tri_d"], g["tri_v"]
    hits = 0
    for i in range(o.shape[0]):
        t = np.zeros(1, np.float32)
        b = np.zeros(3, np.float32)
        h = L.om_intersect_triangle(P(np.ascontiguousarray(o[i])), P(np.ascontiguousarray(d[i])),
                                    P(np.ascontiguousarray(tri[i])), P(t), P(b))
        assert h == g["tri_hit"][i], i
        if h:
            hits += 1
            assert t[0].tobytes() == g["tri_t"][i].tobytes(), i
            assert b.tobytes() == g["tri_bary"][i].tobytes(), i
    assert hits > o.shape[0] // 3


def test_slab_test_matches_reference(oracle, g):
    L = oracle.lib()
    for i in range(g["box_o"].shape[0]):
        o = np.ascontiguousarray(g["box_o"][i])
        inv = (np.float32(1.0) / g["box_d"][i]).astype(np.float32)
        te = np.zeros(1, np.float32)
        h = L.om_intersect_aabb(P(o), P(inv), P(np.ascontiguousarray(g["box_v"][i])),
                                float(g["box_range"][i, 0]), float(g["box_range"][i, 1]), P(te))
        assert h == g["box_hit"][i], i
        if h:
            assert te[0] == g["box_t"][i, 0], i


def test_nudge_bit_exact(oracle, g):
    L = oracle.lib()
    for i in range(g["nudge_p"].shape[0]):
        out = np.zeros(3, np.float32)
        L.om_nudge(P(np.ascontiguousarray(g["nudge_p"][i])), P(np.ascontiguousarray(g["nudge_n"][i])), P(out))
        assert out.tobytes() == g["nudge_out"][i].tobytes(), i


def test_rotation_between_z_bit_exact(oracle, g):
    L = oracle.lib()
    for i in range(g["rotz_n"].shape[0]):
        q = np.zeros(4, np.float32)
        r = np.zeros(3, np.float32)
        L.om_rot_between_z(P(np.ascontiguousarray(g["rotz_n"][i])), P(q))
        L.om_quat_rotate(P(q), P(np.ascontiguousarray(g["rotz_v"][i])), P(r))
        assert q.tobytes() == g["rotz_q"][i].tobytes(), i
        assert r.tobytes() == g["rotz_out"][i].tobytes(), i


def test_hemisphere_uniform_sampling(oracle, g):
    L = oracle.lib()
    for i in range(g["hemi_xi"].shape[0]):
        d = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        L.om_hemi_uniform(float(g["hemi_xi"][i, 0]), float(g["hemi_xi"][i, 1]), P(d), P(pdf))
        assert pdf[0] == g["hemi_pdf"][i]
        np.testing.assert_allclose(d, g["hemi_dir"][i], atol=1e-6)


def test_reflect_normal_aabb_centroid_bit_exact(oracle, g):
    L = oracle.lib()
    for i in range(g["reflect_wi"].shape[0]):
        r = np.zeros(3, np.float32)
        L.om_reflect(P(np.ascontiguousarray(g["reflect_wi"][i])), P(np.ascontiguousarray(g["reflect_n"][i])), P(r))
        assert r.tobytes() == g["reflect_out"][i].tobytes(), i
        n = np.zeros(3, np.float32)
        L.om_triangle_normal(P(np.ascontiguousarray(g["tri_v"][i])), P(n))
        assert n.tobytes() == g["trinorm_out"][i].tobytes(), i
        c = np.zeros(3, np.float32)
        L.om_aabb_centroid(P(np.ascontiguousarray(g["triaabb_out"][i])), P(c))
        assert c.tobytes() == g["centroid_out"][i].tobytes(), i
    aabb = oracle.tri_aabbs(g["tri_v"])
    assert aabb.tobytes() == g["triaabb_out"].tobytes()


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libmrhref.so")),
                    reason="reference host build (oracle/_ref) not present in this checkout")
def test_oracle_against_live_reference_build(oracle):
    """Fresh random cases against the reference build itself, where it exists."""
    L, R = oracle.lib(), oracle.ref()
    rng = np.random.default_rng(7)
    for _ in range(3000):
        q = rng.integers(0, 1 << 21, 3)
        assert L.om_morton_compose63(*map(int, q)) == R.ref_compose63(*map(int, q))
        o = rng.uniform(-1, 1, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        t1, t2 = np.zeros(1, np.float32), np.zeros(1, np.float32)
        b1, b2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
        h1 = L.om_intersect_triangle(P(o), P(d), P(tri), P(t1), P(b1))
        h2 = R.ref_intersect_triangle(P(o), P(d), P(tri), P(t2), P(b2))
        assert h1 == h2
        if h1:
            assert t1.tobytes() == t2.tobytes() and b1.tobytes() == b2.tobytes()
