"""Pins the oracle's random streams and pinhole camera.

The GPU renders are compared with the oracle sample for sample, so a wrong
PCG constant, a mirrored image axis or a fov on the wrong axis in the oracle
would pass every render parity test.  These checks tie the oracle to
published known answers and to the geometry of a pinhole camera instead.

- PCG32: O'Neill's pcg32-demo output for pcg32_srandom_r(42, 54)
  (pcg-c-basic, "pcg32_random_r"), the first six words.
- SplitMix64: the published stream of Vigna's splitmix64.c from state 0
  (each output adds the golden gamma to the state, then finalises).
- Camera (GPUCameraPinhole.cuh:140-240): the centre ray looks along
  normalize(gaze - pos); the left/right edge rays subtend fov_x/2 on either
  side; the bottom/top edge rays subtend fov_y/2 with
  fov_y = 2 atan(tan(fov_x / 2) * H / W); image +x runs along
  normalize(cross(gaze - pos, up)); row 0 is the bottom row.
"""
import math

import numpy as np
import pytest


# ------------------------------------------------------------------ RNG
def test_pcg32_known_answer(oracle):
    got = oracle.pcg32_stream(42, 54, 6)
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert [int(x) for x in got] == want


def test_splitmix64_known_answer(oracle):
    gamma = 0x9E3779B97F4A7C15
    want = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f, 0xf88bb8a8724c81ec]
    # om_splitmix64(z) = finaliser(z + gamma): the k-th output of the
    # generator started at state 0 is om_splitmix64((k - 1) * gamma)
    got = [oracle.splitmix64((k * gamma) & (2 ** 64 - 1)) for k in range(4)]
    assert got == want


def test_path_uniforms_are_top_24_bits(oracle):
    """A path's uniforms are the top 24 bits of its PCG32 stream, which is
    keyed by (SplitMix64((seed << 32) ^ id), id)."""
    seed, sid = 7, 123456
    key = oracle.splitmix64(((seed << 32) ^ sid) & (2 ** 64 - 1))
    words = oracle.pcg32_stream(key, sid, 8)
    u = oracle.path_uniforms(seed, sid, 8)
    assert np.array_equal(u, (words >> 8).astype(np.float32) * np.float32(2.0 ** -24))
    assert (u >= 0).all() and (u < 1).all()


# --------------------------------------------------------------- camera
W, H = 7, 5
FOV_X = math.radians(60.0)
POS = np.array([1.0, 2.0, 3.0])
GAZE = np.array([0.5, -1.0, 0.2])
UP = np.array([0.1, 1.0, 0.3])      # not orthogonal to the view direction


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _angle(a, b):
    return math.acos(max(-1.0, min(1.0, float(np.dot(_unit(a), _unit(b))))))


@pytest.fixture(scope="module")
def cam(oracle):
    return oracle.camera(POS, GAZE, UP, FOV_X, 0.1, 100.0, W, H)


def _ray(oracle, cam, px, py, jx, jy):
    o, d = oracle.camera_ray_at(cam, W, H, 1, py * W + px, jx, jy)
    assert np.allclose(o, POS, atol=0)
    return d.astype(np.float64)


def _basis():
    fwd = _unit(GAZE - POS)
    right = _unit(np.cross(GAZE - POS, UP))
    up = _unit(np.cross(right, fwd))
    return fwd, right, up


def test_camera_centre_ray_looks_at_gaze(oracle, cam):
    fwd, _, _ = _basis()
    d = _ray(oracle, cam, W // 2, H // 2, 0.5, 0.5)
    assert np.allclose(_unit(d), fwd, atol=2e-6)


def test_camera_horizontal_fov(oracle, cam):
    fwd, right, _ = _basis()
    left = _ray(oracle, cam, 0, H // 2, 0.0, 0.5)
    rite = _ray(oracle, cam, W - 1, H // 2, 1.0, 0.5)
    assert abs(_angle(left, fwd) - FOV_X / 2) < 2e-6
    assert abs(_angle(rite, fwd) - FOV_X / 2) < 2e-6
    assert abs(_angle(left, rite) - FOV_X) < 4e-6
    # image +x runs along normalize(cross(gaze - pos, up))
    assert np.dot(left, right) < 0 < np.dot(rite, right)
    xs = [np.dot(_ray(oracle, cam, x, H // 2, 0.5, 0.5), right) for x in range(W)]
    assert all(a < b for a, b in zip(xs, xs[1:]))


def test_camera_vertical_fov_and_bottom_row_first(oracle, cam):
    fwd, _, up = _basis()
    fov_y = 2.0 * math.atan(math.tan(FOV_X / 2) * H / W)
    bottom = _ray(oracle, cam, W // 2, 0, 0.5, 0.0)
    top = _ray(oracle, cam, W // 2, H - 1, 0.5, 1.0)
    assert abs(_angle(bottom, fwd) - fov_y / 2) < 2e-6
    assert abs(_angle(top, fwd) - fov_y / 2) < 2e-6
    # row 0 is the bottom row: its rays point below the view direction
    assert np.dot(bottom, up) < 0 < np.dot(top, up)
    ys = [np.dot(_ray(oracle, cam, W // 2, y, 0.5, 0.5), up) for y in range(H)]
    assert all(a < b for a, b in zip(ys, ys[1:]))


def test_camera_up_is_orthogonalised(oracle, cam):
    """The vertical edge rays lie in the plane of the view direction and the
    given up vector (the basis is re-orthogonalised, not taken as given)."""
    fwd, right, _ = _basis()
    for py, jy in ((0, 0.0), (H - 1, 1.0)):
        d = _ray(oracle, cam, W // 2, py, 0.5, jy)
        assert abs(np.dot(_unit(d), right)) < 2e-6


def test_camera_strata_tile_the_pixel(oracle, cam):
    """With S x S strata the jittered sample of stratum (sx, sy) at jitter
    (jx, jy) is the single-stratum ray at ((sx + jx) / S, (sy + jy) / S)."""
    S = 3
    c3 = oracle.camera(POS, GAZE, UP, FOV_X, 0.1, 100.0, W, H)
    px, py = 2, 3
    for sy in range(S):
        for sx in range(S):
            sid = (py * W + px) * S * S + sy * S + sx
            _, d = oracle.camera_ray_at(c3, W, H, S, sid, 0.25, 0.75)
            ref = _ray(oracle, cam, px, py, (sx + 0.25) / S, (sy + 0.75) / S)
            assert np.allclose(_unit(d), _unit(ref), atol=2e-6)


def test_camera_ray_draws_jitter_from_the_stream(oracle, cam):
    """om_camera_ray uses the path stream's first two uniforms as jitter."""
    seed, sid = 3, (2 * W + 4)
    u = oracle.path_uniforms(seed, sid, 2)
    _, d, _ = oracle.camera_ray(cam, W, H, 1, seed, sid)
    _, d2 = oracle.camera_ray_at(cam, W, H, 1, sid, float(u[0]), float(u[1]))
    assert np.array_equal(d, d2)
