"""Launch configuration: the autotuner picks a configuration that the context then uses, and every
configuration (LDS staging on / off, persistent blocks per CU) gives exactly the same result."""
import numpy as np
import pytest

from sbrsolver_amd import Context, make_params, scenes

pytestmark = pytest.mark.gpu


def test_autotune_and_configurations_do_not_change_results():
    cfg = scenes.config2(tess=5)
    with Context(0) as ctx:
        ctx.scene_from_soup(*cfg["soup"])
        ctx.materials(cfg["materials"])
        p = make_params(**dict(cfg["trace"], receptionModel=1))
        base = ctx.trace(p, cfg["rx"])
        f0, n0 = base.fields, base.paths()["prx"].copy()
        print("[case] autotune", flush=True)
        t = ctx.autotune(p, cfg["rx"])
        assert t == ctx.get_tuning() and t["lds_scene"] in (0, 1) and t["blocks_per_cu"] >= 1
        for lds, bpc in [(0, 1), (1, 2), (0, 3), (-1, 0), (t["lds_scene"], t["blocks_per_cu"])]:
            ctx.set_tuning(lds, bpc)
            r = ctx.trace(p, cfg["rx"])
            assert np.array_equal(r.fields, f0, equal_nan=True), (lds, bpc)
            assert np.array_equal(r.paths()["prx"], n0), (lds, bpc)
