"""The guiding octree's triangle voxeliser on the GPU (mrh_svo_voxelize)
against the oracle (om_voxelize): every record, in order, bit for bit, for
random triangle soups of mixed sizes and orientations (with degenerate and
axis-aligned ones), the Cornell box (whole-wall triangles, up to a million
pixels each) and the displaced sphere grid of the frame benchmark at full
size (1,966,080 triangles, depth 10); the octree built from the records
against the oracle's; the size query and the argument checks."""
import numpy as np
import pytest

from meturay_amd import native as N
from test_gpu_guide import assert_same_svo, dev, gpu_view, host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def cube_of(oracle, tris, depth):
    p = tris.reshape(-1, 3)
    return oracle.svo_cube(p.min(0), p.max(0), depth)


def voxelize_gpu(ctx, tris, cube, depth, prims=False):
    d_t = dev(np.ascontiguousarray(tris, np.float32).reshape(-1))
    torch.cuda.synchronize()
    n = len(tris)
    count = ctx.svo_voxelize(d_t, n, cube, depth)
    d_c = torch.full((max(count, 1),), -1, dtype=torch.int32, device="cuda")
    d_p = torch.full((max(count, 1),), -1, dtype=torch.int32, device="cuda") if prims else None
    torch.cuda.synchronize()
    again = ctx.svo_voxelize(d_t, n, cube, depth, d_c, count, d_prims=d_p)
    ctx.sync()
    assert again == count
    if prims:
        return host(d_c, np.uint32)[:count], host(d_p, np.uint32)[:count]
    return host(d_c, np.uint32)[:count]


def soup(rng, n):
    c = rng.uniform(-0.9, 0.9, (n, 1, 3))
    s = rng.choice([0.003, 0.02, 0.2, 0.7], (n, 1, 1))
    t = (c + s * rng.normal(size=(n, 3, 3))).clip(-1, 1).astype(np.float32)
    t[::17, :, 1] = t[::17, :1, 1]                 # axis-aligned planes
    t[5::29, 2] = t[5::29, 1]                      # degenerate (two equal corners)
    return t.reshape(n, 9)


@pytest.mark.parametrize("depth", [4, 7, 10])
def test_matches_oracle_on_triangle_soup(ctx, oracle, depth):
    rng = np.random.default_rng(depth)
    tris = soup(rng, 3000)
    cube = cube_of(oracle, tris, depth)
    print(f"[vox] soup depth {depth}", flush=True)
    got, gprim = voxelize_gpu(ctx, tris, cube, depth, prims=True)
    want, wprim = oracle.voxelize(tris, cube, depth, prims=True)
    assert len(got) == len(want) and np.array_equal(got, want)
    assert np.array_equal(gprim, wprim)
    assert len(want) > 3000


@pytest.mark.parametrize("depth", [8, 10])
def test_matches_oracle_on_cornell_box(ctx, oracle, depth):
    tris = np.ascontiguousarray(N.gen_cornell()["tris"], np.float32).reshape(-1, 9)
    cube = cube_of(oracle, tris, depth)
    got = voxelize_gpu(ctx, tris, cube, depth)
    want = oracle.voxelize(tris, cube, depth)
    assert np.array_equal(got, want)
    # the octree over the records, as the oracle builds it
    svo = N.Svo(ctx)
    d = dev(got)
    torch.cuda.synchronize()
    svo.build(d.data_ptr(), len(got), cube.tolist(), depth)
    o = oracle.Svo.build(want, cube.tolist(), depth)
    assert_same_svo(gpu_view(ctx, svo), o, filtered=False)


def test_full_size_sphere_grid(ctx, oracle):
    sc = N.gen_sphere_grid(64, 16, 16, 2)
    tris = np.ascontiguousarray(sc["tris"], np.float32).reshape(-1, 9)
    depth = 10
    cube = cube_of(oracle, tris, depth)
    print(f"[vox] sphere grid {len(tris)} triangles depth {depth}", flush=True)
    got = voxelize_gpu(ctx, tris, cube, depth)
    want = oracle.voxelize(tris, cube, depth)
    assert len(got) == len(want) and np.array_equal(got, want)
    assert (got < (1 << 30)).all()


def test_size_query_and_capacity(ctx, oracle):
    rng = np.random.default_rng(2)
    tris = soup(rng, 200)
    cube = cube_of(oracle, tris, 6)
    d_t = dev(tris.reshape(-1))
    torch.cuda.synchronize()
    count = ctx.svo_voxelize(d_t, 200, cube, 6)
    assert count == len(oracle.voxelize(tris, cube, 6))
    small = torch.full((count,), 7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert ctx.svo_voxelize(d_t, 200, cube, 6, small, count - 1) == count   # too small: nothing written
    ctx.sync()
    assert (small.cpu().numpy() == 7).all()
    assert ctx.svo_voxelize(d_t, 0, cube, 6) == 0


def test_rejects_bad_arguments(ctx):
    d_t = torch.zeros(9, device="cuda")
    torch.cuda.synchronize()
    for cube, depth in (([0, 0, 0, 1], 0), ([0, 0, 0, 1], 11), ([0, 0, 0, 0], 5), ([0, 0, float("nan"), 1], 5)):
        with pytest.raises(N.MrhError):
            ctx.svo_voxelize(d_t, 1, cube, depth)
