"""lymph_add_point_sources_noalloc (staging in the library's file-scope device
array instead of per-call device buffers) gives the same RHS as
lymph_add_point_sources, bit for bit, on two calls in a row -- the second with
fewer sources, so stale staging from the first call would show -- and refuses
more sources than the array holds.  Called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import _lib, mesh as M
from lymph3d_apsc_amd._lib import lib, ptr

pytestmark = pytest.mark.gpu

MAX_SOURCES = 16384  # the staging array's capacity (include/lymph_gpu.h)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def _sources(m, n, seed):
    rng = np.random.default_rng(seed)
    pts = np.ascontiguousarray(rng.uniform(0.05, 0.95, (n, 3)))
    tets = np.ascontiguousarray(m.locate(pts), np.int32)
    f = np.ascontiguousarray(rng.standard_normal((n, 3)))
    return pts, tets, f


def _rhs_after(ctx, fn, calls):
    """Zero RHS slot, then `calls` applied with entry point `fn`; the slot."""
    ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_DYN)
    outs = []
    for pts, tets, f in calls:
        rc = fn(ctx.h, pts.shape[0], ptr(pts), ptr(tets), ptr(f), _lib.SLOT_RHS_DYN)
        assert rc == 0, lib().lymph_last_error().decode()
        outs.append(ctx.get_slot(_lib.SLOT_RHS_DYN))
    return outs


@pytest.mark.parametrize("name,p", [("cube3agg", 2), ("grown", 3)])
def test_noalloc_point_sources_match_allocating_entry_point(name, p):
    print(f"[no-alloc] {name} p={p}", flush=True)
    if name == "grown":
        m = M.cube_mesh(4, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=3)
    else:
        m = M.cube_mesh(3, agglomerate=True)
    calls = [_sources(m, 700, 1), _sources(m, 37, 2)]  # second call smaller
    with L.Context(m, p) as ctx:
        ctx.make_matrices_free()
        ref = _rhs_after(ctx, lib().lymph_add_point_sources, calls)
        got = _rhs_after(ctx, lib().lymph_add_point_sources_noalloc, calls)
        assert np.abs(ref[0]).max() > 0
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        # the second call's own contribution, as the difference of the two states
        assert np.array_equal(got[1] - got[0], ref[1] - ref[0])
        # beyond the staging array: a clear error, the slot unchanged
        big = _sources(m, MAX_SOURCES + 1, 3)
        before = ctx.get_slot(_lib.SLOT_RHS_DYN)
        rc = lib().lymph_add_point_sources_noalloc(ctx.h, big[0].shape[0], ptr(big[0]), ptr(big[1]),
                                                   ptr(big[2]), _lib.SLOT_RHS_DYN)
        assert rc == _lib.ERR_ARG
        assert b"at most 16384" in lib().lymph_last_error()
        assert np.array_equal(ctx.get_slot(_lib.SLOT_RHS_DYN), before)
        # zero sources is a no-op
        z = np.zeros((0, 3))
        assert lib().lymph_add_point_sources_noalloc(ctx.h, 0, ptr(z), None, ptr(z), _lib.SLOT_RHS_DYN) == 0


def test_noalloc_full_capacity():
    """Exactly as many sources as the staging array holds."""
    print("[no-alloc] full capacity", flush=True)
    m = M.cube_mesh(3, agglomerate=True)
    calls = [_sources(m, MAX_SOURCES, 4), _sources(m, 5, 5)]
    with L.Context(m, 1) as ctx:
        ctx.make_matrices_free()
        ref = _rhs_after(ctx, lib().lymph_add_point_sources, calls)
        got = _rhs_after(ctx, lib().lymph_add_point_sources_noalloc, calls)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
