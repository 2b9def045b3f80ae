"""Long misses (longMiss = 1, an extension), in the oracle, CPU.  The reference ends a ray that
misses the scene after a bounce with a 100 m stub (RTSolver.cpp:224-256), so a receiver farther
along it cannot capture it.  Closed form on a flat ground (two triangles, 4 km square) with the
transmitter at height h and receivers at height z and distance x: the ground-reflected path
reflects at x h / (h + z) and its last segment is (x z / (h + z)) sqrt(1 + ((h + z) / x)^2) long.
Receivers whose last segment exceeds 100 m keep only the direct path with stubs, and gain the
ground path, reflecting at the closed-form point, with long misses; nearer receivers are the same
either way."""
import numpy as np

from oracle import oracle as O

H, Z = 10.0, 10.0


def ground_soup():
    L = 2000.0
    c = np.array([[-L, -L, 0, L, -L, 0, L, L, 0], [-L, -L, 0, L, L, 0, -L, L, 0]], np.float32)
    n = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    return c.reshape(-1), n.reshape(-1)


def ground():
    return O.Scene.from_soup(*ground_soup())


def _paths(res):
    out = {}
    for k in range(len(res.prx)):
        o, m = int(res.poff[k]), int(res.pnseg[k])
        out.setdefault(int(res.prx[k]), []).append((m, res.ssrc[o:o + m].copy(), res.stgt[o:o + m].copy()))
    return out


def test_long_miss_against_two_ray_geometry():
    sc = ground()
    xs = np.array([40.0, 80.0, 150.0, 260.0, 400.0, 700.0])
    rx = np.array([[x, 0.0, Z] for x in xs])
    mats = np.array([[1.8e9, 5.0, 0.0, 5.0, 0.0, 0.01]])
    kw = dict(tx=(0.0, 0.0, H), tess=7, maxR=1, maxT=0, frequency=1.8e9, receptionModel=1)
    a = _paths(O.trace(sc, O.make_params(**kw), rx, mats, nthreads=8))
    b = _paths(O.trace(sc, O.make_params(**dict(kw, longMiss=1)), rx, mats, nthreads=8))
    for r, x in enumerate(xs):
        last = x * Z / (H + Z) * np.sqrt(1 + ((H + Z) / x) ** 2)
        refl_b = [p for p in b.get(r, []) if p[0] == 2]
        refl_a = [p for p in a.get(r, []) if p[0] == 2]
        assert len(refl_b) == 1, (x, b.get(r))
        q = refl_b[0][2][0]   # target of the first segment: the reflection point
        assert np.allclose(q, [x * H / (H + Z), 0.0, 0.0], atol=1e-9 * x)
        assert any(p[0] == 1 for p in b[r]) and any(p[0] == 1 for p in a[r])   # direct path
        if last > 100.0:
            assert refl_a == []
        else:
            assert len(refl_a) == 1 and np.array_equal(refl_a[0][2], refl_b[0][2])
