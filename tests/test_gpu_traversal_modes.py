"""The three closest-hit traversals (the wide tree with its winner check, the reference tree in the
reference's order, the reference tree with the clipping cuts) give the same interaction trees --
every bounce's hit triangle, hit point and path length -- and the same kept paths and fields, on
the full headline workload (urban canyon, level 9, 16.3 M bounces) and on the dense indoor scene
(207k triangles, leaf BVHs) at level 6."""
import numpy as np
import pytest

from sbrsolver_amd import Context, make_params, scenes

pytestmark = pytest.mark.gpu


def _run(ctx, cfg, mode, **over):
    kw = dict(cfg["trace"], exactTraversal=mode, **over)
    r = ctx.trace(make_params(**kw), cfg["rx"])
    t = r.trees()
    keep = {k: t[k] for k in ("launch", "depth", "code", "tri", "pl", "tgt")}
    return keep, r.fields.copy(), r.paths()


def _same(a, b, label):
    ta, fa, pa = a
    tb, fb, pb = b
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), f"{label}: tree field {k}"
    assert np.array_equal(fa, fb, equal_nan=True), f"{label}: fields"
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), f"{label}: path field {k}"


@pytest.mark.parametrize("workload", ["urban_L9", "indoor_L6"])
def test_traversal_modes_agree_at_size(workload):
    if workload == "urban_L9":
        cfg = scenes.config2(tess=9)
        over = dict(receptionModel=1)
    else:
        cfg = scenes.config3(tess=6, rx_grid=(40, 20))
        over = dict(receptionModel=1)
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        print(f"[case] {workload} wide tree", flush=True)
        a = _run(ctx, cfg, 0, **over)
        n = len(a[0]["tri"])
        for mode in (2, 1):
            print(f"[case] {workload} traversal mode {mode} ({n} bounces)", flush=True)
            _same(a, _run(ctx, cfg, mode, **over), f"{workload} mode {mode}")
