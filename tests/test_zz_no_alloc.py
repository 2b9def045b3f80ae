"""The allocation-free closest-hit entry point (sbr_closest_hit_noalloc) against sbr_closest_hit,
both through the C-ABI: two calls in a row, the second smaller (the staging array keeps the first
call's contents), and a call above the array's ray limit, which is refused before any GPU work."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from sbrsolver_amd import lib

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SBR_E_LIMIT = -6
MAX_RAYS = 399442   # SBR_NOALLOC_MAX_RAYS


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    org = lo + (hi - lo) * (rng.random((n, 3)) * 1.4 - 0.2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(org), np.ascontiguousarray(d)


def _call(fn, h, org, d, exact):
    n = len(org)
    tri = np.full(n, -7, np.int32)
    pt = np.full((n, 3), np.nan)
    dist = np.full(n, np.nan)
    rc = fn(h, org.ctypes.data, d.ctypes.data, n, int(exact), tri.ctypes.data, pt.ctypes.data, dist.ctypes.data)
    return rc, tri, pt, dist


def test_noalloc_closest_hit_equals_allocating_entry_point(gpu_ctx):
    L = lib()
    assert L.sbr_closest_hit_noalloc is not None
    for scene in ("bahen.stl", "office.stl"):
        print("[case] noalloc", scene, flush=True)
        gpu_ctx.scene_from_stl(os.path.join(GOLDEN, "stl", scene))
        tr = gpu_ctx.scene_triangles()
        v = tr["verts"].reshape(-1, 3)
        lo, hi = v.min(0), v.max(0)
        for n, exact, seed in ((300000, 0, 1), (1000, 1, 2), (MAX_RAYS, 0, 3), (17, 0, 4)):
            org, d = _rays(n, seed, lo, hi)
            rc0, t0, p0, d0 = _call(L.sbr_closest_hit, gpu_ctx.h, org, d, exact)
            rc1, t1, p1, d1 = _call(L.sbr_closest_hit_noalloc, gpu_ctx.h, org, d, exact)
            assert rc0 == 0 and rc1 == 0, (rc0, rc1)
            assert np.array_equal(t0, t1), (scene, n)
            assert np.array_equal(p0, p1, equal_nan=True) and np.array_equal(d0, d1, equal_nan=True), (scene, n)
            assert np.any(t1 >= 0) and np.any(t1 < 0)
    org, d = _rays(MAX_RAYS + 1, 5, lo, hi)
    rc, t, _, _ = _call(L.sbr_closest_hit_noalloc, gpu_ctx.h, org, d, 0)
    assert rc == SBR_E_LIMIT and np.all(t == -7)
    assert b"exceed" in L.sbr_last_error(gpu_ctx.h)
    sys.stdout.flush()
