"""Occlusion test of corrected paths (checkOcclusion = 1) on the GPU: the same kept paths and fields
as the oracle on the office scene, with the reference's mesh indexing and with the physical one."""
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params
from tests.parity import compare
from tests.office import KW, MATS, STL
from tests.office import scene_and_rx as _rx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("over", [dict(), dict(q1MeshIndexing=0, tess=4, maxT=2, typedDedup=1, wallWidth=0.1)],
                         ids=["L3_reference_indexing", "L4_physical"])
def test_occlusion_matches_oracle(gpu_ctx, over):
    print("[case] office occlusion", over, flush=True)
    sc, rx = _rx()
    gpu_ctx.scene_from_stl(STL)
    gpu_ctx.materials(MATS)
    kw = dict(KW, checkOcclusion=1, **over)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
    compare(o, g, f"office occlusion {over}")
