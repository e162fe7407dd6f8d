"""Pins the oracle's 4-wide box test (om_slab_box_test: planes ordered by
the direction's sign, distances fma(plane, inv, -o * inv), directions below
2^-64 in magnitude clamped to +-2^-64) by geometry computed independently
in float64, and against the reference's own slab test on the golden boxes.

A planted mistake (near and far planes swapped, the sign test reversed, a
dropped tmin/tmax, or -o*inv with the wrong sign) fails these."""
import ctypes as C
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def slab(oracle, o, d, box, tmin, tmax):
    te = np.zeros(1, np.float32)
    h = oracle.lib().om_slab_box_test(P(np.ascontiguousarray(o, np.float32)), P(np.ascontiguousarray(d, np.float32)),
                                      P(np.ascontiguousarray(box, np.float32)), float(tmin), float(tmax), P(te))
    return bool(h), float(te[0])


def exact_interval(o, d, box, tmin, tmax):
    """[entry, exit] of the ray segment in the box, float64; None if empty."""
    lo, hi = float(tmin), float(tmax)
    for a in range(3):
        if d[a] == 0.0:
            if not (box[a] <= o[a] <= box[3 + a]):
                return None
            continue
        t0 = (float(box[a]) - float(o[a])) / float(d[a])
        t1 = (float(box[3 + a]) - float(o[a])) / float(d[a])
        lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
    return (lo, hi) if lo <= hi else None


def test_slab_box_matches_float64_geometry(oracle):
    rng = np.random.default_rng(7)
    n_hit = 0
    for _ in range(20000):
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        c = rng.uniform(-2, 2, 3)
        e = rng.uniform(0.01, 1.0, 3)
        box = np.concatenate([c - e, c + e]).astype(np.float32)
        tmin = np.float32(rng.choice([0.0, 1e-4, 0.5]))
        tmax = np.float32(rng.choice([1e30, 2.0, 4.0]))
        h, te = slab(oracle, o, d, box, tmin, tmax)
        iv = exact_interval(o, d, box, tmin, tmax)
        if iv is None:
            # a miss in float64 may be a hit in float only within rounding of the boundary
            if h:
                assert te <= tmax
            continue
        lo, hi = iv
        width = hi - lo
        if width > 1e-4 * max(1.0, abs(lo), abs(hi)):
            assert h, (o, d, box, tmin, tmax)
            n_hit += 1
            assert abs(te - lo) <= 1e-5 * max(1.0, abs(lo)), (te, lo)
    assert n_hit > 500


def test_slab_box_empty_slots_and_axis_parallel_rays(oracle):
    empty = np.array([np.inf] * 3 + [-np.inf] * 3, np.float32)
    rng = np.random.default_rng(8)
    for _ in range(2000):
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        for tmax in (1e30, np.inf):
            assert not slab(oracle, o, d, empty, 0.0, tmax)[0]
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    # directions with zero components (either sign): the origin's slab decides
    for dz in (0.0, -0.0):
        d = np.array([1.0, 0.0, dz], np.float32)
        assert slab(oracle, [-1, 0.5, 0.5], d, box, 0, 1e30)[0]
        assert not slab(oracle, [-1, 0.5, 1.5], d, box, 0, 1e30)[0]
        assert not slab(oracle, [-1, -0.5, 0.5], d, box, 0, 1e30)[0]
        h, te = slab(oracle, [-1, 0.5, 0.5], d, box, 0, 1e30)
        assert te == 1.0
        # behind the origin, or beyond tmax
        assert not slab(oracle, [2, 0.5, 0.5], d, box, 0, 1e30)[0]
        assert not slab(oracle, [-1, 0.5, 0.5], d, box, 0, 0.5)[0]
    # negative directions: the max plane is the near one
    h, te = slab(oracle, [3, 0.5, 0.5], [-1, 0, 0], box, 0, 1e30)
    assert h and te == 2.0
    h, te = slab(oracle, [0.5, 0.5, 0.5], [-1, 0, 0], box, 0.25, 1e30)
    assert h and te == 0.25                        # inside: the entry is tmin


def test_slab_box_agrees_with_reference_slab_on_goldens(oracle):
    """The golden boxes of the reference's own IntersectsAABB: the ordered
    test returns the same verdict wherever the entry and exit are not within
    rounding of each other, and entry distances within 1e-5."""
    g = np.load(GOLD, allow_pickle=False)
    agree = 0
    for i in range(g["box_o"].shape[0]):
        o, d, box = g["box_o"][i], g["box_d"][i], g["box_v"][i]
        tmin, tmax = g["box_range"][i]
        h, te = slab(oracle, o, d, box, tmin, tmax)
        iv = exact_interval(o.astype(np.float64), d.astype(np.float64), box, tmin, tmax)
        near_tie = iv is not None and (iv[1] - iv[0]) <= 1e-5 * max(1.0, abs(iv[0]))
        if near_tie:
            continue
        assert h == bool(g["box_hit"][i]), i
        if h:
            assert abs(te - g["box_t"][i, 0]) <= 1e-5 * max(1.0, abs(g["box_t"][i, 0])), i
        agree += 1
    assert agree > 0.95 * g["box_o"].shape[0]
