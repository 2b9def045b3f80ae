"""Serving loop on the GPU: queries submitted while others run (several slots, two submitting
threads) give exactly the fields of the same query traced alone, and of sbr_trace on a plain
context, and the fields of the oracle's trace of the same transmitter."""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, Server, make_params, scenes
from tests.parity import worst_field_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def urban_ctx():
    cfg = scenes.config2(tess=4)
    ctx = Context(0)
    ctx.scene_from_soup(*cfg["soup"])
    ctx.materials(cfg["materials"])
    yield ctx, cfg
    ctx.close()


def test_results_identical_alone_and_in_the_mix(urban_ctx):
    ctx, cfg = urban_ctx
    _, boxes, _ = scenes.urban_scene()
    qs = scenes.coverage_queries(14, boxes, tess=(3, 6), max_rx=3000)
    qs[5] = (dict(qs[5][0], maxT=1), qs[5][1])                      # one query with transmission
    qs[7] = (dict(qs[7][0], receptionModel=0), qs[7][1])            # one with the reference sphere
    print("[case] serving: 14 queries, 3 slots", flush=True)
    with Server(ctx, slots=3) as srv:
        # query 0 alone
        q0 = srv.submit(make_params(**qs[0][0]), qs[0][1])
        rc, info = srv.poll(q0, wait_ms=-1)
        assert rc == 0, srv.error(q0)
        alone = srv.fields(q0, len(qs[0][1]))
        srv.release(q0)
        # all queries, query 0 again among them, from two threads
        ids = {}
        lock = threading.Lock()

        def client(part):
            for k in part:
                q = srv.submit(make_params(**qs[k][0]), qs[k][1])
                with lock:
                    ids[k] = q

        t1 = threading.Thread(target=client, args=(range(0, 14, 2),))
        t2 = threading.Thread(target=client, args=(range(1, 14, 2),))
        t1.start(); t2.start(); t1.join(); t2.join()
        mixed, slots = {}, set()
        for k, q in ids.items():
            rc, info = srv.poll(q, wait_ms=-1)
            assert rc == 0, srv.error(q)
            slots.add(info["slot"])
            mixed[k] = srv.fields(q, len(qs[k][1]))
            srv.release(q)
        st = srv.stats()
    assert np.array_equal(mixed[0], alone, equal_nan=True)
    assert len(slots) > 1, "queries did not spread over slots"
    assert st["completed"] == 15 and st["failed"] == 0 and st["p50_ms"] <= st["p99_ms"]
    for k, (kw, rx) in enumerate(qs):
        ref = ctx.trace(make_params(**kw), rx).fields
        assert np.array_equal(mixed[k], ref, equal_nan=True), f"query {k}"


def test_served_fields_match_oracle(urban_ctx):
    """Served queries (three at once on three slots: a capture-cell query, one with the reference
    sphere, one with transmission) against the FP64 oracle's trace of the same transmitter."""
    ctx, cfg = urban_ctx
    _, boxes, _ = scenes.urban_scene()
    rng = np.random.default_rng(7)
    qs = []
    for kw, _ in scenes.coverage_queries(3, boxes, seed=7, tess=(5, 5), max_rx=1):
        c = np.asarray(kw["tx"])     # 120 receivers in the 200 m square around the transmitter
        qs.append((kw, np.column_stack([c[0] + rng.uniform(-100, 100, 120), c[1] + rng.uniform(-100, 100, 120),
                                        np.full(120, 1.5)])))
    qs[1] = (dict(qs[1][0], receptionModel=0), qs[1][1])
    qs[2] = (dict(qs[2][0], maxT=1), qs[2][1])
    sc = O.Scene.from_soup(*cfg["soup"])
    print("[case] serving against the oracle: 3 queries, level 5", flush=True)
    with Server(ctx, slots=3) as srv:
        ids = [srv.submit(make_params(**kw), rx) for kw, rx in qs]
        served = []
        for q, (kw, rx) in zip(ids, qs):
            rc, _ = srv.poll(q, wait_ms=-1)
            assert rc == 0, srv.error(q)
            served.append(srv.fields(q, len(rx)))
            srv.release(q)
    for k, ((kw, rx), f) in enumerate(zip(qs, served)):
        o = O.trace(sc, O.make_params(**kw), rx, cfg["materials"], nthreads=8)
        assert np.nanmax(np.abs(o.fields)) > 0, f"query {k}: nothing received"
        worst = worst_field_error(f, o.fields, f"served query {k}")
        print(f"[case] served query {k}: {len(rx)} receivers, worst field rel err {worst:.2e}", flush=True)
