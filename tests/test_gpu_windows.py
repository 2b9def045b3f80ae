"""Launch windows: a trace whose interaction trees are processed in many windows of root items
(small window_nodes) gives exactly the single-window result -- fields, kept paths in order, path
fields, bounce count -- for one transmitter and for a batch."""
import numpy as np
import pytest

from sbrsolver_amd import Context, make_params, scenes

pytestmark = pytest.mark.gpu


def _all(ctx, res):
    return dict(fields=res.fields, paths=res.paths(), pf=res.path_fields(), bounces=res.bounces,
                windows=res.stats["windows"])


@pytest.mark.parametrize("window_nodes", [4096, 30_000])
def test_windows_equal_single_window(window_nodes):
    cfg = scenes.config2(tess=5)
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        for model in (0, 1):
            p = make_params(**dict(cfg["trace"], receptionModel=model))
            print(f"[case] windows {window_nodes} model {model}", flush=True)
            ctx.set_tuning(window_nodes=0)
            a = _all(ctx, ctx.trace(p, cfg["rx"]))
            ctx.set_tuning(window_nodes=window_nodes)
            b = _all(ctx, ctx.trace(p, cfg["rx"]))
            assert a["windows"] == 1 and b["windows"] >= 2
            assert a["bounces"] == b["bounces"]
            assert np.array_equal(a["fields"], b["fields"], equal_nan=True)
            for k in a["paths"]:
                assert np.array_equal(a["paths"][k], b["paths"][k]), k
            assert np.array_equal(a["pf"], b["pf"], equal_nan=True)


def test_batch_windows_equal_single_window():
    cfg = scenes.config4(n_tx=5, rx_n=40)
    plist = [make_params(**dict(t, tess=min(t["tess"], 5))) for t in cfg["txs"]]
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        a = ctx.trace_batch(plist, cfg["rx"])
        fa, ta, pa = a.fields, a.path_tx(), a.paths()    # results are read before the next trace
        ctx.set_tuning(window_nodes=30_000)
        b = ctx.trace_batch(plist, cfg["rx"])
        assert b.stats["windows"] > 3
        assert np.array_equal(fa, b.fields, equal_nan=True)
        assert np.array_equal(ta, b.path_tx())
        pb = b.paths()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k


def test_first_trace_of_a_context_in_windows():
    """A context whose first trace runs in several launch windows (the path export of each window
    reads the trace's own parameters) equals a single-window trace on another context."""
    cfg = scenes.config4(n_tx=3, rx_n=30)
    plist = [make_params(**dict(t, tess=min(t["tess"], 5))) for t in cfg["txs"]]
    with Context(0) as ref:
        ref.scene_from_soup(*cfg["soup"])
        ref.materials(cfg["materials"])
        a = ref.trace_batch(plist, cfg["rx"])
        fa, pa = a.fields, a.paths()
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        ctx.set_tuning(window_nodes=30_000)
        b = ctx.trace_batch(plist, cfg["rx"])
        assert b.stats["windows"] >= 2
        assert np.array_equal(fa, b.fields, equal_nan=True)
        pb = b.paths()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k
