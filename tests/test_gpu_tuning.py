"""Launch configuration: the autotuner picks a configuration that the context then uses, and every
configuration (LDS staging on / off, persistent blocks per CU, traversal stack entries in LDS with
the rest in device memory, grid blocks per resident block) gives exactly the same result."""
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
        assert 0 <= t["stack_lds"] <= 64 and 0 <= t["oversub"] <= 64
        for lds, bpc, sl, os_ in [(0, 1, 0, 0), (1, 2, 0, 0), (0, 3, 0, 0), (-1, 0, 0, 0), (1, 0, 1, 0), (0, 0, 2, 1),
                                  (-1, 0, 0, 16), (t["lds_scene"], t["blocks_per_cu"], t["stack_lds"], t["oversub"])]:
            print(f"[case] lds {lds} blocks/CU {bpc} stack in LDS {sl} oversub {os_}", flush=True)
            ctx.set_tuning(lds, bpc, 0, sl, os_)
            r = ctx.trace(p, cfg["rx"])
            assert np.array_equal(r.fields, f0, equal_nan=True), (lds, bpc, sl, os_)
            assert np.array_equal(r.paths()["prx"], n0), (lds, bpc, sl, os_)
