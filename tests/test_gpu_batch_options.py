"""Per-transmitter options in one batch: a batch that mixes the reference's behaviour, the physical
mode (slab walls, typed dedup, occlusion test, long misses, s/p transport), the tabulated pattern
and the reference sphere gives each transmitter exactly its individual trace's paths and fields
(the options live in the transmitter records, and the occlusion variant of the dedup pass runs for
the whole batch while testing only the transmitters that ask for it).  Mesh indexing is a property of
the batch, not of a transmitter."""
import numpy as np
import pytest

from sbrsolver_amd import make_params, scenes
from tests.test_pattern import F_PH, F_TH, table_from

pytestmark = pytest.mark.gpu


def test_mixed_option_batch_equals_individual_traces(gpu_ctx):
    print("[case] urban L4: batch of 4 transmitters with different options", flush=True)
    cfg = scenes.config4(n_tx=4, rx_n=40)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    gpu_ctx.set_tx_pattern(table_from(F_TH, F_PH, 19, 36))
    # mesh indexing is one per batch (the library rejects a mix): physical indexing for all four
    base = [dict(t, tess=min(t["tess"], 5), maxR=min(t["maxR"], 4), q1MeshIndexing=0) for t in cfg["txs"]]
    plist = [make_params(**base[0]),
             make_params(**dict(base[1], physical=True, wallWidth=0.2, maxT=1)),
             make_params(**dict(base[2], antennaType=3)),
             make_params(**dict(base[3], receptionModel=0, typedDedup=1))]
    b = gpu_ctx.trace_batch(plist, cfg["rx"])
    fb, tb, pb, pfb = b.fields.copy(), b.path_tx(), b.paths(), b.path_fields()   # read before the next trace
    assert fb.shape[0] == len(plist)
    for t, p in enumerate(plist):
        s = gpu_ctx.trace(p, cfg["rx"])
        assert np.array_equal(fb[t], s.fields, equal_nan=True), t
        sp = s.paths()
        sel = np.flatnonzero(tb == t)
        assert len(sel) == len(sp["prx"]) and len(sel) > 0, t
        for k in ("prx", "plaunch", "pnseg"):
            assert np.array_equal(pb[k][sel], sp[k]), (t, k)
        assert np.array_equal(pfb[sel], s.path_fields(), equal_nan=True), t
