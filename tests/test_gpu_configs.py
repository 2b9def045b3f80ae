"""The batched and served configurations against the oracle, not against the library itself.

* Batched multi-transmitter run (the 256-transmitter urban configuration): the first 8 of its
  transmitters with icosahedron level 6 or 7 (4-6 reflections each), traced in one sbr_trace_batch
  call over 400 of the shared receivers; each transmitter's kept paths, corrected vertices, per-path
  and per-receiver fields equal the oracle's trace of that transmitter alone.
* Serving stream on the seeded 1M-triangle, 2000-building city: small coverage queries submitted
  to a 3-slot server at once (so they overlap on the device) give the oracle's fields; the same
  queries traced directly give the oracle's paths.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Server, make_params, scenes
from tests.parity import worst_field_error, compare

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


class _Slice:
    """One transmitter's part of a batched result, shaped like a single trace's Result."""

    def __init__(self, res, t, bounces):
        p = res.paths()
        sel = np.flatnonzero(res.path_tx() == t)
        self._paths = {k: p[k][sel] for k in ("prx", "plaunch", "pnseg")}
        segs = np.concatenate([np.arange(p["poff"][k], p["poff"][k] + p["pnseg"][k]) for k in sel]) \
            if len(sel) else np.zeros(0, np.int64)
        self._paths["poff"] = np.concatenate([[0], np.cumsum(self._paths["pnseg"])[:-1]]) if len(sel) \
            else np.zeros(0, np.int64)
        for k in ("stri", "scop", "stype", "ssrc", "stgt"):
            self._paths[k] = p[k][segs]
        self._pf = res.path_fields()[sel]
        self.fields = res.fields[t]
        self.bounces = bounces

    def paths(self):
        return self._paths

    def path_fields(self):
        return self._pf


def test_batched_transmitters_match_oracle(gpu_ctx):
    cfg = scenes.config4()
    txs = [t for t in cfg["txs"] if t["tess"] <= 7][:8]
    assert len(txs) == 8 and {t["tess"] for t in txs} <= {6, 7}
    rx = cfg["rx"][3::25]
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    say("batched: 8 transmitters, levels", [t["tess"] for t in txs], "receivers", len(rx))
    res = gpu_ctx.trace_batch([make_params(**t) for t in txs], rx)
    stats = res.stats
    sc = O.Scene.from_soup(*cfg["soup"])
    total = 0
    for i, t in enumerate(txs):
        o = O.trace(sc, O.make_params(**t), rx, cfg["materials"], nthreads=16)
        total += o.bounces
        worst = compare(o, _Slice(res, i, o.bounces), f"batched tx {i}")
        say("batched tx", i, "level", t["tess"], "R", t["maxR"], "paths", o.np, "worst", worst)
    assert stats["bounces"] == total


@pytest.fixture(scope="module")
def city():
    say("city scene")
    name, soup, boxes, half, mats = scenes.config5_scene()
    assert len(soup[2]) > 900_000
    return name, soup, boxes, half, mats


def test_city_serving_matches_oracle(gpu_ctx, city):
    name, soup, boxes, half, mats = city
    gpu_ctx.scene_from_soup(*soup)
    gpu_ctx.materials(mats)
    say(name, "uploaded")
    # transmitters of the serving stream's kind (levels 5-6), each with 150 receivers in the 300 m square around it
    rng = np.random.default_rng(55)
    qs = []
    for kw, _ in scenes.coverage_queries(6, boxes, seed=20261203 + 55, half=0.5 * half, tess=(5, 6), max_rx=1):
        c = np.asarray(kw["tx"])
        rx = np.column_stack([c[0] + rng.uniform(-150, 150, 150), c[1] + rng.uniform(-150, 150, 150), np.full(150, 1.5)])
        qs.append((kw, rx))
    sc = O.Scene.from_soup(*soup)
    say("oracle scene built")
    with Server(gpu_ctx, slots=3) as srv:
        ids = [srv.submit(make_params(**kw), rx) for kw, rx in qs]    # all at once: they overlap
        served = []
        for q, (kw, rx) in zip(ids, qs):
            rc, info = srv.poll(q, wait_ms=-1)
            assert rc == 0, srv.error(q)
            served.append(srv.fields(q, len(rx)))
            srv.release(q)
    npaths = []
    for k, ((kw, rx), f) in enumerate(zip(qs, served)):
        o = O.trace(sc, O.make_params(**kw), rx, mats, nthreads=16)
        worst_field_error(f, o.fields, f"city query {k} (served)")
        g = gpu_ctx.trace(make_params(**kw), rx)
        compare(o, g, f"city query {k} (traced)")
        say("city query", k, "level", kw["tess"], "receivers", len(rx), "paths", o.np, "bounces", o.bounces)
        npaths.append(o.np)
    assert sum(npaths) > 100
