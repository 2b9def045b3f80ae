"""Duplicate-removal table sizing: a trace whose tables fill up restarts with tables four times
larger and gives exactly the result of a trace that started large enough; a single-transmitter trace
after a batch of 256 starts from tables scaled down with it and gives the result of the same trace
on a fresh context."""
import numpy as np
import pytest

from sbrsolver_amd import Context, make_params, scenes

pytestmark = pytest.mark.gpu


def _snap(r):
    """Fields, kept paths and stats of a result (a context's next trace replaces them)."""
    return r.fields.copy(), r.paths(), r.stats


def _same(a, b, label):
    assert np.array_equal(a[0], b[0], equal_nan=True), label
    for k in a[1]:
        assert np.array_equal(a[1][k], b[1][k]), (label, k)


@pytest.mark.parametrize("model", [1, 0])
def test_table_restart_equals_large_table(model):
    cfg = scenes.config2(tess=6)
    p = make_params(**dict(cfg["trace"], receptionModel=model))
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        print(f"[case] urban L6 model {model}: automatic tables", flush=True)
        big = _snap(ctx.trace(p, cfg["rx"]))
        sb = big[2]
        print(f"[case] urban L6 model {model}: 1024-entry tables", flush=True)
        ctx.set_tuning(table_entries=1024)
        small = _snap(ctx.trace(p, cfg["rx"]))
        ss = small[2]
        print("[case] table entries", sb["table_entries"], ss["table_entries"], "restarts", ss["table_restarts"],
              flush=True)
        assert ss["table_restarts"] >= 1 and ss["table_entries"] == 1024 * 4 ** ss["table_restarts"]
        assert sb["table_restarts"] == 0 and ss["winners"] == sb["winners"] > 0
        _same(small, big, f"model {model}")


def test_single_after_batch_starts_smaller():
    cfg = scenes.config4()
    plist = [make_params(**t) for t in cfg["txs"]]
    with Context(0) as fresh:
        fresh.scene_from_soup(*cfg["soup"])
        fresh.materials(cfg["materials"])
        print("[case] single transmitter on a fresh context", flush=True)
        ref = _snap(fresh.trace(plist[0], cfg["rx"]))
        ref_entries = ref[2]["table_entries"]
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        print("[case] batch of 256 transmitters, then the first alone", flush=True)
        bst = ctx.trace_batch(plist, cfg["rx"]).stats
        one = _snap(ctx.trace(plist[0], cfg["rx"]))
        print("[case] table entries: batch", bst["table_entries"], "single after it", one[2]["table_entries"],
              "fresh single", ref_entries, flush=True)
        assert ref_entries <= one[2]["table_entries"] < bst["table_entries"]
        _same(one, ref, "single after batch")
