"""Rays aimed at the edges and corners of axis-aligned boxes, whose faces are flat leaves of the
reference tree: for about 1 % of them the reference's FP64 slab test rejects the leaf of the closest
triangle that the ray hits (a line through a box edge, where the slab endpoints of two axes agree only
up to rounding), so the reference returns a farther triangle or nothing.  The wide traversal finds
that closer triangle first; its check on the reference tree fails and the ray is redone there.  All
three traversal modes must give the oracle's reference result on every ray, the redone ones (those
where the oracle's brute force over all triangles finds a closer hit than its BVH traversal)
included."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, scenes

pytestmark = pytest.mark.gpu

BOXES = [(0, 0, 0, 1, 1, 1), (2.5, 0.25, 0, 3.75, 1.5, 0.8), (-3, -2, 0, -1.2, -0.7, 2.3), (0.3, 3, 0.5, 2.1, 4.4, 1.9),
         (-2.2, 2.0, 0.1, -0.6, 3.3, 0.7), (1.7, -2.9, 0.2, 3.1, -1.3, 2.6)]


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    org, dirs = np.zeros((n, 3)), np.zeros((n, 3))
    for i in range(n):
        b = BOXES[i % len(BOXES)]
        lo, hi = np.array(b[:3], float), np.array(b[3:], float)
        p = lo + (hi - lo) * rng.random(3)   # a point on an edge (two coordinates at bounds) or a corner
        ax = rng.permutation(3)[:2]
        for a in ax:
            p[a] = lo[a] if rng.random() < 0.5 else hi[a]
        if i % 3 == 0:
            a = [k for k in range(3) if k not in ax][0]
            p[a] = lo[a] if rng.random() < 0.5 else hi[a]
        o = p + rng.normal(size=3) * rng.uniform(0.5, 6.0)
        d = p - o
        org[i], dirs[i] = o, d / np.linalg.norm(d)
    return org, dirs


def test_wide_winner_check_redoes_grazing_rays():
    s = scenes.Soup()
    for b in BOXES:
        s.box(*b)
    soup = s.arrays()
    sc = O.Scene.from_soup(*soup)
    org, d = _rays(20000, 5)
    ref = [sc.closest_hit(org[i], d[i]) for i in range(len(d))]
    redone = 0
    for i in range(len(d)):
        tb, _, db = sc.closest_hit(org[i], d[i], brute=True)
        if tb >= 0 and (ref[i][0] < 0 or db < ref[i][2]):
            redone += 1
    print(f"[case] {redone} of {len(d)} rays: the reference rejects the leaf of a closer hit", flush=True)
    assert redone > 100
    with Context(0) as ctx:
        ctx.scene_from_soup(*soup)
        for mode in (0, 1, 2):
            print(f"[case] closest hit, traversal mode {mode}", flush=True)
            tri, pt, dist = ctx.closest_hit(org, d, exact=mode)
            for i in range(len(d)):
                ot, op, od = ref[i]
                assert tri[i] == ot, (mode, i, tri[i], ot)
                if ot >= 0:
                    assert np.array_equal(pt[i], op) and dist[i] == od, (mode, i)
