"""Dense indoor scene (about 207k triangles, tessellated floor and ceiling that the reference SAH
leaves as leaves of thousands of triangles): closest hits of the leaf-BVH search (through the wide
tree, and through the reference tree with its clipping cuts) equal the exact traversal and the oracle, including rays aimed at shared vertices and edges of the tessellation
(ties broken by DFS position); a small trace equals the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, make_params, scenes
from tests.parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def indoor():
    print("[case] indoor scene build", flush=True)
    cfg = scenes.config3(tess=4, maxR=2, maxT=2, maxTotal=3, rx_grid=(40, 20))
    ctx = Context(0)
    ctx.scene_from_soup(*cfg["soup"])
    ctx.materials(cfg["materials"])
    sc = O.Scene.from_soup(*cfg["soup"])
    yield cfg, ctx, sc
    ctx.close()


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    org = np.column_stack([rng.uniform(0.5, 49.5, n), rng.uniform(0.5, 49.5, n), rng.uniform(0.2, 3.8, n)])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # every third ray aims at a floor or ceiling grid vertex, every third at a grid edge midpoint
    k = np.arange(n)
    tgt = np.column_stack([rng.integers(1, 199, n) * 0.25, rng.integers(1, 199, n) * 0.25,
                           np.where(rng.random(n) < 0.5, 0.0, 4.0)])
    tgt[k % 3 == 2, 0] += 0.125
    aim = k % 3 != 0
    v = tgt[aim] - org[aim]
    d[aim] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return org, d


def test_indoor_closest_hit_leaf_bvh_exact(indoor):
    cfg, ctx, sc = indoor
    org, d = _rays(3000, 11)
    print("[case] indoor closest hit", flush=True)
    t1, p1, d1 = ctx.closest_hit(org, d, exact=0)
    t0, p0, d0 = ctx.closest_hit(org, d, exact=1)
    t2, p2, d2 = ctx.closest_hit(org, d, exact=2)
    assert np.array_equal(t1, t0) and np.array_equal(p1, p0) and np.array_equal(d1, d0)
    assert np.array_equal(t2, t0) and np.array_equal(p2, p0) and np.array_equal(d2, d0)
    for i in range(0, len(d), 3):
        ot, op, od = sc.closest_hit(org[i], d[i])
        assert t1[i] == ot, (i, t1[i], ot)
        if ot >= 0:
            assert np.array_equal(p1[i], op) and d1[i] == od


@pytest.mark.parametrize("model", [0, 1])
def test_indoor_trace_matches_oracle(indoor, model):
    cfg, ctx, sc = indoor
    print(f"[case] indoor trace model {model}", flush=True)
    kw = dict(cfg["trace"], receptionModel=model)
    g = ctx.trace(make_params(**kw), cfg["rx"])
    o = O.trace(sc, O.make_params(**kw), cfg["rx"], cfg["materials"], nthreads=8)
    compare(o, g, f"indoor_L4_R2T2max3_model{model}")
