"""C-ABI surface checks that need no GPU: the library loads, exports every
function include/mrh.h declares, host-only generators are deterministic, and
argument validation fails with a status code instead of crashing."""
import ctypes as C

import numpy as np
import pytest

from meturay_amd import native as N


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    declared = N.header_symbols()
    assert len(declared) >= 25
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert set(declared) <= set(N._SIGS)


def test_generators_are_deterministic():
    a = N.gen_random_tris(1000, 1)
    b = N.gen_random_tris(1000, 1)
    assert a.tobytes() == b.tobytes()
    c = a.reshape(-1, 3, 3)
    centre = c.mean(axis=1)
    assert np.all(np.abs(c - centre[:, None, :]) <= 0.0200001)
    r = N.gen_random_rays(5000, 2)
    assert np.allclose(np.linalg.norm(r[3:6], axis=0), 1.0, atol=1e-5)
    assert np.all((r[:3] >= 0) & (r[:3] < 1))


def test_scene_generators_shapes():
    sc = N.gen_cornell(4)
    assert sc["tris"].shape == (12 + 2 + 20 * 4 ** 4, 9)
    batches = sc["prim_key"] >> 24
    assert (batches == N.BATCH_EMISSIVE).sum() == 2
    g = N.gen_sphere_grid(64, 16, 16, 2)
    assert g["tris"].shape[0] == 64 * 64 * 2 * 16 * 15
    kinds = np.bincount(g["prim_key"] >> 24, minlength=5)
    assert kinds[0] == 0 and all(kinds[1:5] > 0)
    em = (kinds[4] / (2 * 16 * 15)) / (64 * 64)
    assert 0.03 < em < 0.07


def test_invalid_arguments_return_status():
    lib = N.lib()
    assert lib.mrh_context_destroy(None) == N.ERR_INVALID
    assert lib.mrh_gen_cornell(9, None) == N.ERR_INVALID
    p = N.RenderParams(1920, 1080, 4, 10, 3, 0, 0, 0)
    nbytes = C.c_uint64(0)
    assert lib.mrh_render_bytes(C.byref(p), C.byref(nbytes)) == N.OK
    assert nbytes.value > 33_000_000 * 128
    h = C.c_void_p()
    rc = lib.mrh_context_create(-1, C.byref(h))
    assert rc != N.OK and not h.value
