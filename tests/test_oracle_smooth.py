"""Pins the oracle's smooth-shading frames (om_pt.c: orthogonal_vector,
space_quat, vertex_frame, q_slerp, q_bary_slerp, make_surface_tbn).

OrthogonalVector and TransformGen::Space are plain float arithmetic and must
equal the reference's host-compiled RayLib bit for bit; SLerp / BarySLerp
evaluate acos and sin with the library's own polynomials and must agree
with the reference (libm) within 2e-6.  Closed forms pin the composition: a
vertex frame maps its normal to +Z, a corner of a smooth triangle shades with
that corner's frame, equal corner normals give that normal everywhere."""
import ctypes as C
import os

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libmrhref.so")
has_ref = pytest.mark.skipif(not os.path.exists(REF), reason="reference host build not present")


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def f32(*a):
    return np.ascontiguousarray(np.array(a, np.float32).reshape(-1))


def unit(rng, k=3):
    v = rng.normal(size=k)
    return (v / np.linalg.norm(v)).astype(np.float32)


def quat_rotate(q, v):
    """rotate v by unit quaternion q = (w, x, y, z), float64"""
    w, x, y, z = (float(c) for c in q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, np.float64)


def frame(oracle, n):
    out = np.zeros(12, np.float32)
    nn = np.ascontiguousarray(np.tile(np.asarray(n, np.float32), 3))
    oracle.lib().om_corner_frames(P(nn), 1, P(out))
    return out[:4]


@has_ref
def test_frames_bit_exact_against_reference(oracle):
    L, R = oracle.lib(), oracle.ref()
    rng = np.random.default_rng(31)
    a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
    qa, qb = np.zeros(4, np.float32), np.zeros(4, np.float32)
    for i in range(4000):
        v = unit(rng) * np.float32(rng.uniform(0.1, 3.0))
        if i % 10 == 0:
            v[rng.integers(0, 3)] = 0.0                      # axis planes
        v = np.ascontiguousarray(v, np.float32)
        L.om_orthogonal_vector(P(v), P(a))
        R.ref_orthogonal_vector(P(v), P(b))
        assert a.tobytes() == b.tobytes(), v
        # an orthonormal basis from the restated tangent steps
        n = unit(rng)
        t = np.cross(n, unit(rng)).astype(np.float32)
        t /= np.linalg.norm(t)
        bb = np.cross(n, t).astype(np.float32)
        x, y, z = (np.ascontiguousarray(w, np.float32) for w in (t, bb, n))
        if i % 7 == 0:
            x = np.ascontiguousarray(-x)                    # left-handed: Space negates x
        L.om_space_quat(P(x), P(y), P(z), P(qa))
        R.ref_space_quat(P(x), P(y), P(z), P(qb))
        assert qa.tobytes() == qb.tobytes(), (x, y, z)


@has_ref
def test_slerp_matches_reference(oracle):
    L, R = oracle.lib(), oracle.ref()
    rng = np.random.default_rng(32)
    got, ref = np.zeros(4, np.float32), np.zeros(4, np.float32)
    worst = 0.0
    for i in range(4000):
        q0, q1, q2 = unit(rng, 4), unit(rng, 4), unit(rng, 4)
        if i % 5 == 0:                                       # nearby frames, as on a smooth mesh
            q1 = (q0 + 0.05 * unit(rng, 4)).astype(np.float32)
            q1 /= np.linalg.norm(q1)
            q2 = (q0 + 0.05 * unit(rng, 4)).astype(np.float32)
            q2 /= np.linalg.norm(q2)
        if i % 11 == 0:
            q1 = q0.copy()                                   # parallel: the lerp branch
        t = np.float32(rng.uniform(0, 1))
        L.om_qslerp(P(q0), P(q1), float(t), P(got))
        R.ref_slerp(P(q0), P(q1), float(t), P(ref))
        worst = max(worst, float(np.abs(got - ref).max()))
        if i % 11 == 0:
            assert got.tobytes() == ref.tobytes()
        a = np.float32(rng.uniform(0, 1))
        b = np.float32(rng.uniform(0, 1 - a))
        if i % 13 == 0:
            a, b = np.float32(0.0), np.float32(0.0)          # the third corner itself
        L.om_bary_slerp(P(q0), P(q1), P(q2), float(a), float(b), P(got))
        R.ref_bary_slerp(P(q0), P(q1), P(q2), float(a), float(b), P(ref))
        worst = max(worst, float(np.abs(got - ref).max()))
    assert worst <= 2e-6, worst


def test_vertex_frame_maps_normal_to_z(oracle):
    rng = np.random.default_rng(33)
    for i in range(3000):
        n = unit(rng)
        if i < 6:
            n = np.zeros(3, np.float32)
            n[i % 3] = 1.0 if i < 3 else -1.0                # the axes
        q = frame(oracle, n)
        assert abs(np.linalg.norm(q.astype(np.float64)) - 1.0) < 1e-6
        # world-to-tangent: n -> +Z, and the tangent X is perpendicular to n
        assert np.allclose(quat_rotate(q, n), [0, 0, 1], atol=2e-6)
        tx = quat_rotate([q[0], -q[1], -q[2], -q[3]], [1, 0, 0])
        assert abs(float(np.dot(tx, n))) < 2e-6


def test_smooth_surface_closed_forms(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(34)
    q, nrm = np.zeros(4, np.float32), np.zeros(3, np.float32)
    for _ in range(500):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        e = np.cross(tri[3:6] - tri[0:3], tri[6:9] - tri[0:3])
        face = (e / np.linalg.norm(e)).astype(np.float32)
        # corner normals near the face normal, on the viewer's side
        ns = [(face + 0.3 * unit(rng)) for _ in range(3)]
        ns = [(v / np.linalg.norm(v)).astype(np.float32) for v in ns]
        tq = np.ascontiguousarray(np.concatenate([frame(oracle, v) for v in ns]), np.float32)
        d = f32(*(-face))                                    # front side
        for k, (ba, bb) in enumerate([(1, 0), (0, 1), (0, 0)]):
            L.om_smooth_frame(P(tri), P(tq), float(ba), float(bb), P(d), P(q), P(nrm))
            assert np.allclose(nrm, ns[k], atol=3e-6), k     # a corner shades with its own frame
            assert np.allclose(q, tq[4 * k:4 * k + 4], atol=3e-6) or np.allclose(q, -tq[4 * k:4 * k + 4], atol=3e-6)
        # equal corner normals: that normal everywhere
        tq3 = np.ascontiguousarray(np.tile(frame(oracle, ns[0]), 3), np.float32)
        ba, bb = rng.uniform(0, 0.5, 2)
        L.om_smooth_frame(P(tri), P(tq3), float(ba), float(bb), P(d), P(q), P(nrm))
        assert np.allclose(nrm, ns[0], atol=3e-6)
        # interior: a unit normal near the corners' mean (slerping whole
        # frames, tangent twist included, does not follow the normals'
        # geodesic, so only a cone bound holds)
        L.om_smooth_frame(P(tri), P(tq), 1 / 3, 1 / 3, P(d), P(q), P(nrm))
        avg = sum(ns) / 3
        avg /= np.linalg.norm(avg)
        assert abs(float(np.linalg.norm(nrm)) - 1.0) < 2e-6
        assert float(np.dot(nrm, avg)) > 0.95
        # the back side flips the frame: the shading normal turns around
        d2 = f32(*face)
        L.om_smooth_frame(P(tri), P(tq3), float(ba), float(bb), P(d2), P(q), P(nrm))
        assert np.allclose(nrm, -ns[0], atol=3e-6)
