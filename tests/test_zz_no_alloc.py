"""mrh_partition_by_key_noalloc: the partition entry point that allocates
nothing.  Its results must equal mrh_partition_by_key's bit for bit, on two
calls in a row (the second smaller, so stale workspace contents would show),
and it must work while every device allocation of its context is made to
fail.  Sizes above the fixed workspace are refused before any launch."""
import numpy as np
import pytest

from meturay_amd import native as N

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_KEYS = 1 << 22


def _keys(n, seed):
    rng = np.random.default_rng(seed)
    batch = rng.integers(0, 6, n).astype(np.uint32)
    ids = rng.integers(0, 1 << 20, n).astype(np.uint32)
    return (batch << 24) | ids


def _partition(c, fn, keys, sort_begin):
    n = keys.size
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    order = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    offs = torch.full((257,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    c.check(fn(c.h, dk.data_ptr(), n, sort_begin, 24, 32, order.data_ptr(), offs.data_ptr()))
    c.sync()
    return order.cpu().numpy().view(np.uint32)[:n].copy(), offs.cpu().numpy().view(np.uint32).copy()


@pytest.mark.parametrize("sort_begin", [24, 8, 0])
def test_noalloc_equals_partition_by_key(sort_begin):
    print(f"[no-alloc] sort_begin={sort_begin}", flush=True)
    c = N.Context(0)
    try:
        L = N.lib()
        # a larger call, then a smaller one on the same (shared) workspace
        for n, seed in [(1_000_003, 1), (70_001, 2), (1, 3)]:
            keys = _keys(n, seed)
            want_o, want_f = _partition(c, L.mrh_partition_by_key, keys, sort_begin)
            got_o, got_f = _partition(c, L.mrh_partition_by_key_noalloc, keys, sort_begin)
            assert np.array_equal(got_o, want_o), (n, sort_begin)
            assert np.array_equal(got_f, want_f), (n, sort_begin)
            ref = np.argsort(-(keys >> sort_begin).astype(np.int64), kind="stable")
            assert np.array_equal(got_o, ref.astype(np.uint32))
    finally:
        c.close()


def test_noalloc_allocates_nothing_and_limits():
    c = N.Context(0)
    try:
        L = N.lib()
        keys = _keys(MAX_KEYS, 5)
        want_o, want_f = _partition(c, L.mrh_partition_by_key, keys, 0)
        # from here on every device allocation of this context fails
        c.tune("fail_alloc_after", 1)
        got_o, got_f = _partition(c, L.mrh_partition_by_key_noalloc, keys, 0)
        assert np.array_equal(got_o, want_o) and np.array_equal(got_f, want_f)
        small = _keys(5000, 6)
        got_o, got_f = _partition(c, L.mrh_partition_by_key_noalloc, small, 0)
        ref = np.argsort(-small.astype(np.int64), kind="stable").astype(np.uint32)
        assert np.array_equal(got_o, ref)
        assert got_f[256] == 5000
        # empty input: offsets all zero
        z = torch.full((257,), -1, dtype=torch.int32, device="cuda")
        o = torch.empty(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        c.check(L.mrh_partition_by_key_noalloc(c.h, 0, 0, 0, 24, 32, o.data_ptr(), z.data_ptr()))
        c.sync()
        assert not z.cpu().numpy().any()
        # beyond the workspace: refused on the host, before any launch
        rc = L.mrh_partition_by_key_noalloc(c.h, 16, MAX_KEYS + 1, 0, 24, 32, o.data_ptr(), z.data_ptr())
        assert rc == N.ERR_INVALID
        assert b"fixed workspace" in L.mrh_last_error(c.h)
        c.tune("fail_alloc_after", 0)
    finally:
        c.close()
