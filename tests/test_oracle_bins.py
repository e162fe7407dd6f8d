"""Pins the oracle's ray binning (om_svo_bin_rays; the reference's
AccumulateRayHitsToSVO + CollapseRayCounts + KCCheckReducedSVOBins +
PartitionRaysWRTCustomData).

The reference's kernels are CUDA and do not build on the host, so the
oracle is pinned by
  * a hand-worked tree whose bins follow from the rules by hand, and
  * a brute-force reading of the definition on random trees: recursion
    over child lists built from the parent links, ray sets instead of
    counts: a node at or below min_level claims the rays of its subtree
    that no marked descendant claimed once they number at least
    bin_ray_count (always at min_level); each ray's bin is the node that
    claimed it, its own start node when none did.
"""
import numpy as np
import pytest

NO = 0xFFFFFFFF


def brute_bins(t, start, min_level, min_rays):
    depth = t.depth
    min_level = min(min_level, depth)
    level = np.zeros(t.nodes, np.int64)
    for l in range(depth + 1):
        level[t.level_offset[l]:t.level_offset[l] + t.level_count[l]] = l
    kids = [[] for _ in range(t.nodes)]
    for v in range(t.nodes):
        if t.parent[v] != NO:
            kids[int(t.parent[v])].append(v)
    rays_at = [[] for _ in range(t.nodes)]
    for i, v in enumerate(start):
        if v != NO:
            rays_at[int(v)].append(i)
    bins = np.full(len(start), NO, np.uint32)

    def claim(v):
        s = list(rays_at[v])
        for c in kids[v]:
            s += claim(c)
        if level[v] >= min_level and (level[v] == min_level or len(s) >= min_rays):
            bins[s] = v
            return []
        return s

    for i in claim(0):                        # start nodes above min_level
        bins[i] = start[i]
    return bins


def partition(bins):
    key = np.where(bins == NO, np.uint64(1 << 40), bins.astype(np.uint64))
    order = np.argsort(key, kind="stable").astype(np.uint32)
    good = bins[order] != NO
    u, first = np.unique(bins[order][good], return_index=True)
    offsets = np.append(first, good.sum()).astype(np.uint32)
    return order, offsets, u.astype(np.uint32)


def test_hand_worked_tree(oracle):
    # depth 2, cells: A = code 0, B = code 1 (both under level-1 node 0),
    # C = code 8 (under level-1 node 1).  Rays: 3 at A, 1 at B, 1 at C and
    # 2 at cell 2, which is not in the tree (they start at level-1 node 0).
    t = oracle.Svo.build(np.array([0, 1, 8], np.uint32), [0, 0, 0, 1], 2)
    lv1 = {int(t.codes[v]): v for v in range(t.level_offset[1], t.level_offset[2])}
    lv2 = {int(t.codes[v]): v for v in range(t.level_offset[2], t.level_offset[3])}
    codes = np.array([0, 0, 0, 1, 8, 2, 2, NO], np.uint32)
    bins, order, offsets, pbins = t.bin_rays(codes, 0, 3)
    A, P0, root = lv2[0], lv1[0], 0
    # A holds 3 >= 3: its own bin; B (1 ray) hands up to P0, which then holds
    # 1 + 2 = 3: a bin; C (1) hands up to P1 (1 < 3), which hands up to the
    # root, marked at min_level 0
    assert list(bins) == [A, A, A, P0, root, P0, P0, NO]
    assert list(order) == sorted(range(7), key=lambda i: (int(bins[i]), i)) + [7]
    assert list(pbins) == sorted({root, P0, A})
    # min_level = 1: C's ray stops at P1 although P1 holds only one
    bins, *_ = t.bin_rays(codes, 1, 3)
    assert bins[4] == lv1[1]
    # min_level beyond the depth is clamped: every leaf is a bin, the two
    # rays of the absent cell keep their start node
    bins, *_ = t.bin_rays(codes, 7, 100)
    assert list(bins[:7]) == [A, A, A, lv2[1], lv2[8], P0, P0]
    t.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_trees_against_the_definition(oracle, seed):
    rng = np.random.default_rng(seed)
    depth = int(rng.integers(2, 7))
    # clustered tree cells and rays, some rays in cells the tree lacks
    centres = rng.random((5, 3))
    pts = (centres[rng.integers(0, 5, 3000)] + 0.08 * rng.normal(size=(3000, 3))).clip(0, 0.999)
    cube = [0, 0, 0, 1]
    t = oracle.Svo.build(oracle.svo_codes(pts.astype(np.float32), cube, depth), cube, depth)
    n = int(rng.integers(500, 8000))
    rp = (centres[rng.integers(0, 5, n)] + 0.1 * rng.normal(size=(n, 3))).clip(0, 0.999).astype(np.float32)
    codes = oracle.svo_codes(rp, cube, depth)
    codes[rng.random(n) < 0.1] = NO
    start = np.array([NO if c == NO else t.lookup([c])[0] for c in codes], np.uint32)
    for min_level in [0, 1, depth // 2, depth]:
        for min_rays in [1, 7, 64, 4096]:
            bins, order, offsets, pbins = t.bin_rays(codes, min_level, min_rays)
            want = brute_bins(t, start, min_level, min_rays)
            assert np.array_equal(bins, want), (min_level, min_rays)
            o, off, pb = partition(want)
            assert np.array_equal(order, o) and np.array_equal(offsets, off) and np.array_equal(pbins, pb)
    t.close()


def test_no_rays_and_all_skipped(oracle):
    t = oracle.Svo.build(np.array([0, 5, 77], np.uint32), [0, 0, 0, 1], 3)
    bins, order, offsets, pbins = t.bin_rays(np.zeros(0, np.uint32), 1, 4)
    assert len(bins) == 0 and list(offsets) == [0] and len(pbins) == 0
    bins, order, offsets, pbins = t.bin_rays(np.full(5, NO, np.uint32), 1, 4)
    assert np.all(bins == NO) and list(order) == list(range(5)) and list(offsets) == [0]
    t.close()
