"""Occlusion test of corrected paths (checkOcclusion = 1, an extension), in the oracle, CPU.  The
reference re-intersects only each path's own surfaces after image correction (RTSolver.cpp:330-373),
so a corrected segment may cross another wall.  With the test, a corrected multi-bounce path is
dropped when the closest hit along a segment lies nearer than the segment's end, and duplicate
removal keeps the first surviving path of each group.  Pinned against a brute-force closest hit over
all triangles (independent of the BVH): every kept path is unobstructed, every path the reference
keeps stays kept iff it is unobstructed, and no new coplanar group appears."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.parity import oracle_paths, path_table
from tests.office import KW, MATS
from tests.office import scene_and_rx as _rx


def _blocked(sc, res, i):
    o, n = int(res.poff[i]), int(res.pnseg[i])
    if n == 1:
        return False   # the direct path has the line-of-sight test of its own
    for s in range(o, o + n):
        d = res.stgt[s] - res.ssrc[s]
        tri, _, dist = sc.closest_hit(res.ssrc[s], d, brute=True)
        if tri >= 0 and dist < np.linalg.norm(d) * (1 - 1e-7):
            return True
    return False


@pytest.mark.parametrize("q1", [1, 0])
def test_occlusion_against_brute_force(q1):
    sc, rx = _rx()
    kw = dict(KW, q1MeshIndexing=q1)
    a = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
    b = O.trace(sc, O.make_params(**dict(kw, checkOcclusion=1)), rx, MATS, nthreads=8)
    ta, tb = path_table(oracle_paths(a)), path_table(oracle_paths(b))
    assert all(not _blocked(sc, b, i) for i in range(len(tb)))
    kept = set(tb)
    blocked = [_blocked(sc, a, i) for i in range(len(ta))]
    assert any(blocked) and not all(blocked)
    for i, row in enumerate(ta):
        assert (row in kept) == (not blocked[i]), row
    groups_a = {(r[0], r[3]) for r in ta}
    assert {(r[0], r[3]) for r in tb} <= groups_a
