"""Duplicate removal keyed by interaction types (typedDedup = 1) on the GPU: the same kept paths and
fields as the oracle on the office scene (interior walls, so reflections and transmissions share
coplanar-id sequences)."""
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params
from tests.parity import compare
from tests.office import KW, MATS, STL
from tests.office import scene_and_rx as _rx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("over", [dict(), dict(tess=4, maxT=2, wallWidth=0.1)], ids=["L3_R2T1", "L4_R2T2_slabs"])
def test_typed_dedup_matches_oracle(gpu_ctx, over):
    print("[case] office typed dedup", over, flush=True)
    sc, rx = _rx()
    gpu_ctx.scene_from_stl(STL)
    gpu_ctx.materials(MATS)
    kw = dict(KW, typedDedup=1, **over)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
    compare(o, g, f"office typed dedup {over}")
