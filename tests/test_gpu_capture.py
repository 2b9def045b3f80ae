"""Capture pass work split: a listed node goes to the wave-per-node kernel when its candidate region
spans more than 12 receiver-grid cells, a property of the node alone.  So the number of such nodes
does not change between repeated traces or with the launch-window size (which changes how the
nodes are numbered and grouped into waves), and neither do the records."""
import numpy as np
import pytest

from sbrsolver_amd import Context, make_params, scenes

pytestmark = pytest.mark.gpu


def test_heavy_split_is_per_node():
    cfg = scenes.config2(tess=6)
    p = make_params(**dict(cfg["trace"], receptionModel=1, maxR=4, maxT=0, countTraversal=1))
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        seen = []
        for wn in (0, 0, 0, 20_000):
            print(f"[case] urban L6 R4T0 model 1 window_nodes {wn}", flush=True)
            ctx.set_tuning(window_nodes=wn)
            r = ctx.trace(p, cfg["rx"])
            s = r.stats
            seen.append((s["heavy_nodes"], s["capture_nodes"], s["capture_tests"], s["capture_records"],
                         r.fields.copy()))
        assert seen[-1][0] > 0
        for h, n, t, rec, f in seen[1:]:
            assert (h, n, t, rec) == seen[0][:4]
            assert np.array_equal(f, seen[0][4], equal_nan=True)
