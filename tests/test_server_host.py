"""Serving loop on the CPU (host-only context, device -1): queue, admission by HBM budget, polling,
release, statistics and error paths.  Queries fail with SBR_E_STATE on a host-only context, which
exercises the failure path of every query without a GPU."""
import threading

import pytest

from sbrsolver_amd import Context, SbrError, Server, SBR_PENDING, make_params, scenes

SBR_E_ARG, SBR_E_STATE, SBR_E_NOMEM = -1, -2, -4


@pytest.fixture()
def host_ctx():
    cfg = scenes.config1()
    ctx = Context(-1)
    ctx.scene_from_soup(*cfg["soup"])
    ctx.materials(cfg["materials"])
    yield ctx, cfg
    ctx.close()


def test_server_needs_scene_and_materials():
    with Context(-1) as ctx:
        with pytest.raises(SbrError, match="no scene"):
            Server(ctx)
        cfg = scenes.config1()
        ctx.scene_from_soup(*cfg["soup"])
        with pytest.raises(SbrError, match="no materials"):
            Server(ctx)


def test_queries_fail_cleanly_without_device(host_ctx):
    ctx, cfg = host_ctx
    with Server(ctx, slots=2) as srv:
        ids = [srv.submit(make_params(**cfg["trace"]), cfg["rx"]) for _ in range(5)]
        assert len(set(ids)) == 5
        for q in ids:
            rc, info = srv.poll(q, wait_ms=-1)
            assert rc == SBR_E_STATE and info["state"] == 3 and info["status"] == SBR_E_STATE
            assert "no device" in srv.error(q)
            with pytest.raises(SbrError):
                srv.fields(q, len(cfg["rx"]))
            srv.release(q)
        with pytest.raises(SbrError):
            srv.poll(ids[0])          # released: unknown id
        st = srv.stats()
        assert st["submitted"] == 5 and st["failed"] == 5 and st["completed"] == 0 and st["running"] == 0


def test_admission_refuses_query_over_budget(host_ctx):
    ctx, cfg = host_ctx
    with Server(ctx, slots=1, hbm_budget_bytes=50_000_000) as srv:
        small = srv.submit(make_params(**dict(cfg["trace"], tess=2)), cfg["rx"])
        with pytest.raises(SbrError, match="exceeds the HBM budget"):
            srv.submit(make_params(**dict(cfg["trace"], tess=9)), cfg["rx"])
        rc, _ = srv.poll(small, wait_ms=-1)
        assert rc == SBR_E_STATE
        st = srv.stats()
        assert st["refused"] == 1 and st["submitted"] == 1 and st["budget_bytes"] == 50_000_000


def test_bad_arguments(host_ctx):
    ctx, cfg = host_ctx
    with Server(ctx, slots=1) as srv:
        for tess in (12, 13, -1):   # the engine accepts levels 0..11: a level-12 query is refused at submit
            with pytest.raises(SbrError, match="tess"):
                srv.submit(make_params(**dict(cfg["trace"], tess=tess)), cfg["rx"])
        assert srv.stats()["submitted"] == 0
        with pytest.raises(SbrError, match="unknown"):
            srv.poll(12345)


def test_concurrent_submitters(host_ctx):
    """Several host threads submit and poll at once; every query is accounted for exactly once."""
    ctx, cfg = host_ctx
    with Server(ctx, slots=3) as srv:
        out = []
        lock = threading.Lock()

        def client(k):
            for _ in range(20):
                q = srv.submit(make_params(**dict(cfg["trace"], tess=1)), cfg["rx"][: 1 + k])
                rc, info = srv.poll(q, wait_ms=-1)
                with lock:
                    out.append((rc, info["n_rx"], k))
                srv.release(q)

        ts = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert len(out) == 80 and all(rc == SBR_E_STATE and n == 1 + k for rc, n, k in out)
        st = srv.stats()
        assert st["submitted"] == 80 and st["failed"] == 80 and st["queued"] == 0


def test_destroy_with_queued_queries(host_ctx):
    ctx, cfg = host_ctx
    srv = Server(ctx, slots=1)
    for _ in range(50):
        srv.submit(make_params(**cfg["trace"]), cfg["rx"])
    srv.close()   # must not hang: queued queries are failed, the running one finishes
    assert srv.h is None
    assert SBR_PENDING == 1
