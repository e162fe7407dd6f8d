"""Cone tracing through the guiding octree on the GPU (mrh_svo_cone_trace)
against the oracle (om_svo_cone_trace): nodes, t and leaf flags bit for bit
over octrees voxelised from triangles (depth 5, 8 and 10) and from points,
for rays inside and outside the cube, axis-parallel rays, rays with a tmin,
apertures from a ray to a wide cone, and level offsets; then the trace the
field stage makes: 64x64 concentric-octahedral directions (the first-bounce
field size) from points inside the voxelised Cornell box at depth 10."""
import math

import numpy as np
import pytest

from meturay_amd import native as N
from test_gpu_guide import dev, host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def build_both(ctx, oracle, codes, cube, depth):
    svo = N.Svo(ctx)
    d = dev(codes)
    torch.cuda.synchronize()
    svo.build(d.data_ptr(), len(codes), cube.tolist(), depth)
    ctx.sync()
    return svo, oracle.Svo.build(codes, cube.tolist(), depth)


def trace_gpu(ctx, svo, rays, aperture, offset):
    n = len(rays)
    d_r = dev(np.ascontiguousarray(rays, np.float32).reshape(-1))
    d_n = torch.empty(n, dtype=torch.int32, device="cuda")
    d_t = torch.empty(n, dtype=torch.float32, device="cuda")
    d_l = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    svo.cone_trace(d_r, n, aperture, offset, d_n, d_t, d_l)
    ctx.sync()
    return host(d_n, np.uint32), d_t.cpu().numpy(), host(d_l, np.uint32)


def random_rays(rng, n, span=1.6):
    o = rng.uniform(-span, span, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    k = len(d[::9])
    d[::9] = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1, 1], (k, 1))
    d[4::13, rng.integers(0, 3)] = 0.0
    tmin = np.where(rng.random(n) < 0.25, rng.uniform(0, 0.4, n), 0.0)
    return np.hstack([o, d, tmin[:, None]]).astype(np.float32)


@pytest.mark.parametrize("depth", [5, 8, 10])
def test_matches_oracle_on_voxelised_triangles(ctx, oracle, depth):
    rng = np.random.default_rng(depth)
    cube = oracle.svo_cube([-1, -1, -1], [1, 1, 1], depth)
    c = rng.uniform(-0.8, 0.8, (300, 1, 3))
    tris = (c + 0.25 * rng.normal(size=(300, 3, 3))).clip(-1, 1).astype(np.float32).reshape(-1, 9)
    codes = oracle.voxelize(tris, cube, depth)
    svo, o = build_both(ctx, oracle, codes, cube, depth)
    rays = random_rays(rng, 20000)
    for aperture, offset in ((0.0, 0), (1e-3, 0), (3e-2, 0), (0.0, 2), (5e-3, 1)):
        print(f"[cone] depth {depth} aperture {aperture} offset {offset}", flush=True)
        gn, gt, gl = trace_gpu(ctx, svo, rays, aperture, offset)
        on, ot, ol = o.cone_trace(rays, aperture, offset)
        assert np.array_equal(gn, on) and np.array_equal(gl, ol)
        assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
        assert (gn != 0xFFFFFFFF).sum() > 1000


def test_matches_oracle_on_point_octree(ctx, oracle):
    rng = np.random.default_rng(41)
    depth = 7
    cube = oracle.svo_cube([-1, -1, -1], [1, 1, 1], depth)
    codes = oracle.svo_codes(rng.uniform(-1, 1, (5000, 3)).astype(np.float32), cube, depth)
    svo, o = build_both(ctx, oracle, codes, cube, depth)
    rays = random_rays(rng, 10000)
    gn, gt, gl = trace_gpu(ctx, svo, rays, 2e-3, 0)
    on, ot, ol = o.cone_trace(rays, 2e-3, 0)
    assert np.array_equal(gn, on) and np.array_equal(gl, ol)
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))


def test_field_directions_in_cornell_box(ctx, oracle):
    """What a bin's field needs: 64x64 octahedral texel directions from a
    point, cone aperture = the texel's solid angle (4 pi / 4096)."""
    depth = 10
    tris = np.ascontiguousarray(N.gen_cornell()["tris"], np.float32).reshape(-1, 9)
    p = tris.reshape(-1, 3)
    cube = oracle.svo_cube(p.min(0), p.max(0), depth)
    codes = oracle.voxelize(tris, cube, depth)
    svo, o = build_both(ctx, oracle, codes, cube, depth)
    u = (np.arange(64) + 0.5) / 64
    uv = np.stack(np.meshgrid(u, u), -1).reshape(-1, 2).astype(np.float32)
    dirs = oracle.octa_to_dirs(uv)
    rng = np.random.default_rng(5)
    lo, hi = p.min(0), p.max(0)
    origins = lo + (hi - lo) * rng.uniform(0.2, 0.8, (8, 3))
    rays = np.concatenate([np.hstack([np.repeat(org[None], len(dirs), 0), dirs, np.zeros((len(dirs), 1))])
                           for org in origins]).astype(np.float32)
    aperture = 4 * math.pi / 4096
    gn, gt, gl = trace_gpu(ctx, svo, rays, aperture, 0)
    on, ot, ol = o.cone_trace(rays, aperture, 0)
    assert np.array_equal(gn, on) and np.array_equal(gl, ol)
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
    assert (gn != 0xFFFFFFFF).mean() > 0.95          # closed box: almost every cone lands


def test_rejects_bad_arguments(ctx, oracle):
    depth = 4
    cube = oracle.svo_cube([0, 0, 0], [1, 1, 1], depth)
    codes = oracle.svo_codes(np.random.default_rng(0).random((50, 3)).astype(np.float32), cube, depth)
    svo, _ = build_both(ctx, oracle, codes, cube, depth)
    r = torch.zeros(7, device="cuda")
    nn = torch.zeros(1, dtype=torch.int32, device="cuda")
    tt = torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    for ap in (-1.0, float("nan"), float("inf")):
        with pytest.raises(N.MrhError):
            svo.cone_trace(r, 1, ap, 0, nn, tt)
    with pytest.raises(N.MrhError):
        svo.cone_trace(r, 1, 0.0, 0, None, tt)
    empty = N.Svo(ctx)
    with pytest.raises(N.MrhError):
        empty.cone_trace(r, 1, 0.0, 0, nn, tt)


@pytest.mark.parametrize("side", [64, 16])
def test_trace_fields_match_oracle_and_feed_sampling(ctx, oracle, side):
    """Cone-traced fields of 24 bins in the voxelised Cornell box (depth 10),
    then one guided sample per ray from them: both against the oracle."""
    depth = 10
    tris = np.ascontiguousarray(N.gen_cornell()["tris"], np.float32).reshape(-1, 9)
    p = tris.reshape(-1, 3)
    cube = oracle.svo_cube(p.min(0), p.max(0), depth)
    codes = oracle.voxelize(tris, cube, depth)
    svo, o = build_both(ctx, oracle, codes, cube, depth)
    rng = np.random.default_rng(side)
    n = o.nodes
    irr = rng.gamma(0.7, 1.5, (n, 2)).astype(np.float32)
    irr[rng.random((n, 2)) < 0.2] = 0.0
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    bins = 24
    lo, hi = p.min(0), p.max(0)
    origins = np.hstack([lo + (hi - lo) * rng.uniform(0.1, 0.9, (bins, 3)),
                         np.full((bins, 1), float(cube[3]) / 1024 * 0.15 + 1e-5)]).astype(np.float32)
    jit = rng.random((bins, 2)).astype(np.float32)
    d_o, d_j, d_i, d_n = dev(origins.reshape(-1)), dev(jit.reshape(-1)), dev(irr.reshape(-1)), dev(nrm.reshape(-1))
    d_f = torch.empty(bins * side * side, device="cuda")
    d_nodes = torch.empty(bins * side * side, dtype=torch.int32, device="cuda")
    d_t = torch.empty(bins * side * side, device="cuda")
    torch.cuda.synchronize()
    svo.trace_fields(side, side, bins, d_o, d_j, d_i, d_n, d_f, boundary=0.5, d_nodes=d_nodes, d_t=d_t)
    ctx.sync()
    of, on, ot = o.trace_fields(side, side, origins, jit, irr, nrm, boundary=0.5)
    assert np.array_equal(host(d_nodes, np.uint32), on.ravel())
    assert np.array_equal(d_t.cpu().numpy().view(np.uint32), ot.ravel().view(np.uint32))
    assert np.array_equal(d_f.cpu().numpy(), of.ravel())
    # the fields drive the guided samples of each bin's rays
    n_rays = 6000
    offsets = np.linspace(0, n_rays, bins + 1).astype(np.uint32)
    ids = rng.permutation(n_rays).astype(np.uint32)
    xi = rng.random((n_rays, 2)).astype(np.float32)
    d_uv = torch.full((n_rays, 2), float("nan"), device="cuda")
    d_pdf = torch.full((n_rays,), float("nan"), device="cuda")
    d_off, d_ids, d_xi = dev(offsets), dev(ids), dev(xi)
    torch.cuda.synchronize()
    ctx.wfpg_sample_fields(d_f, side, side, bins, d_off, d_ids, n_rays, d_xi, d_uv, d_pdf)
    ctx.sync()
    want = oracle.wfpg_sample_fields(of, offsets, ids, xi, alpha=1.0, ray_count=n_rays)
    assert np.array_equal(d_uv.cpu().numpy(), want["uv"]) and np.array_equal(d_pdf.cpu().numpy(), want["pdf"])


def surface_vertices(rng, tris, n):
    """points on random triangles, random unit wo, luminance; plus a few in
    empty space and outside the cube"""
    t = tris[rng.integers(0, len(tris), n)].reshape(n, 3, 3).astype(np.float64)
    w = rng.dirichlet([1, 1, 1], n)
    p = (w[:, :, None] * t).sum(1)
    p[: n // 20] += rng.normal(scale=0.05, size=(n // 20, 3))
    p[n // 20: n // 10] = rng.uniform(-10, 10, (n // 10 - n // 20, 3))
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    lum = rng.gamma(0.8, 1.0, n)
    return np.hstack([p, wo, lum[:, None]]).astype(np.float32)


def test_payload_matches_oracle_and_drives_the_fields(ctx, oracle):
    """Payload of the voxelised Cornell box (depth 10) from 200k path
    vertices, then the fields it gives and the samples from them: every stage
    against the oracle, bit for bit."""
    depth = 10
    tris = np.ascontiguousarray(N.gen_cornell()["tris"], np.float32).reshape(-1, 9)
    p = tris.reshape(-1, 3)
    cube = oracle.svo_cube(p.min(0), p.max(0), depth)
    codes, prims = oracle.voxelize(tris, cube, depth, prims=True)
    svo, o = build_both(ctx, oracle, codes, cube, depth)
    rng = np.random.default_rng(77)
    verts = surface_vertices(rng, tris, 200000)
    n = o.nodes
    d_t, d_c, d_p, d_v = dev(tris.reshape(-1)), dev(codes), dev(prims), dev(verts.reshape(-1))
    d_irr = torch.empty(2 * n, device="cuda")
    d_nrm = torch.empty(3 * n, device="cuda")
    d_len = torch.empty(n, device="cuda")
    torch.cuda.synchronize()
    svo.payload(d_t, d_c, d_p, len(codes), d_v, len(verts), d_irr, d_nrm, d_len)
    ctx.sync()
    oi, on = o.payload(tris, codes, prims, verts)
    assert np.array_equal(d_irr.cpu().numpy().view(np.uint32), oi.ravel().view(np.uint32))
    assert np.array_equal(d_nrm.cpu().numpy().view(np.uint32), on[:, :3].ravel().view(np.uint32))
    assert np.array_equal(d_len.cpu().numpy().view(np.uint32), on[:, 3].view(np.uint32))
    assert (oi[o.level_offset[depth]:] != 65504.0).mean() > 0.5
    # a vertex mask (0xFFFFFFFF = skipped) is the payload of the kept vertices
    keep = rng.random(len(verts)) < 0.6
    mask = np.where(keep, 0, 0xFFFFFFFF).astype(np.uint32)
    d_irr2 = torch.empty(2 * n, device="cuda")
    d_m = dev(mask)
    torch.cuda.synchronize()
    svo.payload(d_t, d_c, d_p, len(codes), d_v, len(verts), d_irr2, d_nrm, d_mask=d_m)
    ctx.sync()
    oi2, _ = o.payload(tris, codes, prims, np.ascontiguousarray(verts[keep]))
    assert np.array_equal(d_irr2.cpu().numpy().view(np.uint32), oi2.ravel().view(np.uint32))
    assert not np.array_equal(oi2, oi)
    # fields from that payload at bins on the surfaces, then samples
    bins, side = 16, 32
    org = verts[:bins, :3] + 0.01 * verts[:bins, 3:6]
    origins = np.hstack([org, np.full((bins, 1), float(cube[3]) / 1024 * 0.15 + 1e-5)]).astype(np.float32)
    jit = rng.random((bins, 2)).astype(np.float32)
    d_f = torch.empty(bins * side * side, device="cuda")
    torch.cuda.synchronize()
    svo.trace_fields(side, side, bins, dev(origins.reshape(-1)), dev(jit.reshape(-1)), d_irr, d_nrm, d_f)
    ctx.sync()
    of, _, _ = o.trace_fields(side, side, origins, jit, oi, on[:, :3])
    assert np.array_equal(d_f.cpu().numpy(), of.ravel())
