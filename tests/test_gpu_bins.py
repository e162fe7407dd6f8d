"""Ray binning by octree node on the GPU (mrh_svo_bin_rays) against the
oracle (om_svo_bin_rays): bins, the partition order, run offsets, run nodes
and the run count, bit for bit, over trees of depth 3 to 10, clustered and
skipped rays, rays in cells the tree lacks, every collapse regime
(bin_ray_count 1 .. 4096, min_level 0 .. depth and beyond), and at the
guided configuration's scale (octree depth 10, one ray per 1080p pixel)."""
import numpy as np
import pytest

from meturay_amd import native as N
from test_gpu_guide import assert_same_svo, dev, gpu_view, host

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

NO = 0xFFFFFFFF


def points(rng, n, centres, spread):
    return (centres[rng.integers(0, len(centres), n)] + spread * rng.normal(size=(n, 3))).clip(-1.99, 1.99)


def run_case(ctx, oracle, depth, n_tree, n_rays, settings, seed):
    rng = np.random.default_rng(seed)
    cube = [-2.0, -2.0, -2.0, 4.0]
    centres = rng.random((9, 3)) * 3 - 1.5
    tree_codes = oracle.svo_codes(points(rng, n_tree, centres, 0.2).astype(np.float32), cube, depth)
    ray_codes = oracle.svo_codes(points(rng, n_rays, centres, 0.25).astype(np.float32), cube, depth)
    ray_codes[rng.random(n_rays) < 0.07] = NO                       # next-event rays, misses
    svo = N.Svo(ctx)
    d_tree = dev(tree_codes)
    torch.cuda.synchronize()
    svo.build(d_tree.data_ptr(), n_tree, cube, depth)
    o = oracle.Svo.build(tree_codes, cube, depth)
    assert_same_svo(gpu_view(ctx, svo), o, filtered=False)
    d_codes = dev(ray_codes)
    d_bins = torch.empty(max(n_rays, 1), dtype=torch.int32, device="cuda")
    d_order = torch.empty_like(d_bins)
    d_off = torch.empty(n_rays + 1, dtype=torch.int32, device="cuda")
    d_pb = torch.empty_like(d_bins)
    d_cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    for min_level, min_rays in settings:
        print(f"[bins] depth {depth} tree {n_tree} rays {n_rays} min_level {min_level} min_rays {min_rays}",
              flush=True)
        d_off.fill_(-1)
        torch.cuda.synchronize()
        svo.bin_rays(d_codes.data_ptr(), n_rays, min_level, min_rays, d_bins.data_ptr(), d_order.data_ptr(),
                     d_off.data_ptr(), d_pb.data_ptr(), d_cnt.data_ptr())
        ctx.sync()
        bins, order, offsets, pbins = o.bin_rays(ray_codes, min_level, min_rays)
        k = int(host(d_cnt, np.uint32)[0])
        assert k == len(pbins)
        assert np.array_equal(host(d_bins, np.uint32)[:n_rays], bins)
        assert np.array_equal(host(d_order, np.uint32)[:n_rays], order)
        assert np.array_equal(host(d_off, np.uint32)[:k + 1], offsets)
        assert np.array_equal(host(d_pb, np.uint32)[:k], pbins)
        # bins only: the same bins without the partition outputs
        svo.bin_rays(d_codes.data_ptr(), n_rays, min_level, min_rays, d_order.data_ptr())
        ctx.sync()
        assert np.array_equal(host(d_order, np.uint32)[:n_rays], bins)
    o.close()
    svo.close()


@pytest.mark.parametrize("depth", [3, 6, 10])
def test_bins_match_oracle(ctx, oracle, depth):
    settings = [(0, 1), (0, 64), (1, 512), (depth // 2, 7), (depth, 4096), (depth + 3, 100), (2, 4096)]
    run_case(ctx, oracle, depth, 40000, 70001, settings, seed=depth)


def test_bins_at_guided_scale(ctx, oracle):
    # octree depth 10 over ~1.5 M cells, one ray per 1080p pixel, BinRayCount 512
    run_case(ctx, oracle, 10, 2_000_000, 1920 * 1080, [(0, 512), (4, 512)], seed=77)


def test_bins_edge_cases(ctx, oracle):
    cube = [0.0, 0.0, 0.0, 1.0]
    svo = N.Svo(ctx)
    d_bins = torch.zeros(8, dtype=torch.int32, device="cuda")
    d_cnt = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    d_off = torch.full((9,), 9, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(N.MrhError):                               # not built
        svo.bin_rays(d_bins.data_ptr(), 1, 0, 4, d_bins.data_ptr())
    tree = np.array([0, 5, 77], np.uint32)
    d_tree = dev(tree)
    torch.cuda.synchronize()
    svo.build(d_tree.data_ptr(), 3, cube, 3)
    for bad in [0, 4097]:
        with pytest.raises(N.MrhError):
            svo.bin_rays(d_tree.data_ptr(), 3, 0, bad, d_bins.data_ptr())
    with pytest.raises(N.MrhError):                               # a partition needs the order
        svo.bin_rays(d_tree.data_ptr(), 3, 0, 4, d_bins.data_ptr(), None, d_off.data_ptr())
    # no rays: an empty partition
    d_order = torch.zeros(8, dtype=torch.int32, device="cuda")
    svo.bin_rays(d_tree.data_ptr(), 0, 1, 4, d_bins.data_ptr(), d_order.data_ptr(), d_off.data_ptr(),
                 d_bins.data_ptr(), d_cnt.data_ptr())
    ctx.sync()
    assert int(host(d_cnt, np.uint32)[0]) == 0 and int(host(d_off, np.uint32)[0]) == 0
    # every ray skipped
    skipped = dev(np.full(5, NO, np.uint32))
    torch.cuda.synchronize()
    d_pb = torch.zeros(8, dtype=torch.int32, device="cuda")
    svo.bin_rays(skipped.data_ptr(), 5, 1, 4, d_bins.data_ptr(), d_order.data_ptr(), d_off.data_ptr(),
                 d_pb.data_ptr(), d_cnt.data_ptr())
    ctx.sync()
    assert np.all(host(d_bins, np.uint32)[:5] == NO)
    assert list(host(d_order, np.uint32)[:5]) == list(range(5))
    assert int(host(d_cnt, np.uint32)[0]) == 0 and int(host(d_off, np.uint32)[0]) == 0
    svo.close()
