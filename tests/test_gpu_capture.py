"""Capture pass work split: a listed node goes to the wave-per-node kernel when its candidate region
spans more than 12 receiver-grid cells (48 with the reference sphere), a property of the node alone.
So the number of such nodes does not change between repeated traces or with the launch-window size
(which changes how the nodes are numbered and grouped into waves), and neither do the records.

Record-buffer and heavy-list overflow: the dedup pass runs before the host sees that the candidate
pass filled a partition of the record buffer or the list of nodes left to the wave-per-node kernel
(65,536 entries on a fresh context); the window's capture is then redone with larger ones over
tables that already hold part of the result.  A fresh context's first trace that overflows both must
equal the same trace once they have grown."""
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


def test_record_overflow_redo_equals_clean_pass():
    """Reference sphere with the occlusion test (every candidate record goes through the dedup pass:
    millions of records at level 8, above the 2 M first sized), first trace of a fresh context
    against the second, whose record buffer kept its grown size."""
    cfg = scenes.config2(tess=8)
    p = make_params(**dict(cfg["trace"], receptionModel=0, checkOcclusion=1))
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        print("[case] urban L8 reference sphere with occlusion: first trace of a context", flush=True)
        a = ctx.trace(p, cfg["rx"])
        fa, pa, ra = a.fields.copy(), a.paths(), a.stats["capture_records"]
        print("[case] the same trace again", flush=True)
        b = ctx.trace(p, cfg["rx"])
        print("[case] records", ra, "corrected", a.stats["corrected"], b.stats["corrected"], flush=True)
        assert ra == b.stats["capture_records"] and ra > 2_500_000
        assert a.stats["heavy_nodes"] == b.stats["heavy_nodes"] > 65536
        # the redo corrected the records of the first pass again: the first trace counts more
        assert a.stats["corrected"] > b.stats["corrected"]
        assert np.array_equal(fa, b.fields, equal_nan=True)
        pb = b.paths()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k
