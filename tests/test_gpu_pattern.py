"""Tabulated transmitter pattern (antennaType 3) on the GPU: the same paths and fields as the oracle's
table lookup (fields within the parity tolerance), the Hertzian-dipole table against the analytic
Hertzian trace, servers taking the table over, and the error for a trace that asks for a table the
context does not have."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, SbrError, make_params, scenes
from tests.parity import compare
from tests.test_pattern import F_PH, F_TH, hertz_table, table_from

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("model", [0, 1])
def test_table_pattern_matches_oracle(gpu_ctx, model):
    say("table pattern vs oracle, urban L3 R6T1, model", model)
    cfg = scenes.config2(tess=3)
    table = table_from(F_TH, F_PH, 37, 72)   # 5-degree grid, phi-dependent, both polarisations
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    gpu_ctx.set_tx_pattern(table)
    tr = dict(cfg["trace"], receptionModel=model, antennaType=3)
    g = gpu_ctx.trace(make_params(**tr), cfg["rx"])
    O.set_pattern(table)
    o = O.trace(O.Scene.from_soup(*cfg["soup"]), O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    worst = compare(o, g, f"table pattern model {model}")
    say("paths", len(o.prx), "worst field rel err", worst)


def test_hertzian_table_against_analytic_trace(gpu_ctx):
    say("Hertzian table vs antennaType 1, urban L5 R6T1")
    cfg = scenes.config2(tess=5)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    gpu_ctx.set_tx_pattern(hertz_table())
    tr = dict(cfg["trace"], receptionModel=1)
    a = gpu_ctx.trace(make_params(**dict(tr, antennaType=1)), cfg["rx"])
    pa, fa = a.paths(), a.path_fields()
    b = gpu_ctx.trace(make_params(**dict(tr, antennaType=3)), cfg["rx"])
    pb, fb = b.paths(), b.path_fields()
    assert len(pa["prx"]) > 1000
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    mag = np.linalg.norm(fa, axis=1)
    err = np.linalg.norm(fb - fa, axis=1)
    assert np.all(err <= 5e-5 * mag + 1e-12 * mag.max())   # 1-degree bilinear interpolation of sin(theta)
    assert err.max() > 0


def test_server_takes_the_table_over(gpu_ctx):
    say("server from a context with a table pattern")
    from sbrsolver_amd import Server
    cfg = scenes.config2(tess=4)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    gpu_ctx.set_tx_pattern(table_from(F_TH, F_PH, 19, 36))
    p = make_params(**dict(cfg["trace"], receptionModel=1, antennaType=3))
    want = gpu_ctx.trace(p, cfg["rx"]).fields.copy()
    with Server(gpu_ctx, slots=2) as srv:
        q = srv.submit(p, cfg["rx"])
        rc, _ = srv.poll(q, wait_ms=-1)
        assert rc == 0, srv.error(q)
        got = srv.fields(q, len(cfg["rx"]))
        srv.release(q)
    assert np.array_equal(got, want)


def test_table_pattern_needs_a_table():
    say("antennaType 3 without a table")
    cfg = scenes.config2(tess=2)
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        with pytest.raises(SbrError):
            ctx.trace(make_params(**dict(cfg["trace"], antennaType=3)), cfg["rx"])
        with pytest.raises((SbrError, ValueError)):
            ctx.set_tx_pattern(np.zeros((1, 4, 4), np.float32))
