"""Slab walls (wallWidth > 0) on the GPU, on the office scene (interior walls: paths through walls
are kept): the same paths and fields as the oracle; slab walls change the field of every path
through a wall (the reference's infinite walls transmit nothing of its theta and phi components);
the paths do not change and the direct paths keep their fields."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params
from tests.office import KW, MATS, STL
from tests.office import scene_and_rx as _rx
from tests.parity import compare

pytestmark = pytest.mark.gpu


def test_slab_walls_match_oracle(gpu_ctx):
    print("[case] office L3 R2T1: 0.15 m slab walls vs oracle", flush=True)
    sc, rx = _rx()
    gpu_ctx.scene_from_stl(STL)
    gpu_ctx.materials(MATS)
    kw = dict(KW, wallWidth=0.15)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
    assert np.any(o.stype == 2)
    compare(o, g, "office slab walls")


def test_slab_walls_carry_transmissions(gpu_ctx):
    print("[case] office L3 R2T1: infinite walls vs 0.15 m slabs", flush=True)
    _, rx = _rx()
    gpu_ctx.scene_from_stl(STL)
    gpu_ctx.materials(MATS)
    a = gpu_ctx.trace(make_params(**KW), rx)
    pa, fa = a.paths(), a.path_fields()
    b = gpu_ctx.trace(make_params(**dict(KW, wallWidth=0.15)), rx)
    pb, fb = b.paths(), b.path_fields()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    n = len(pa["prx"])
    through = np.array([np.any(pa["stype"][int(pa["poff"][i]):int(pa["poff"][i]) + int(pa["pnseg"][i])] == 2)
                        for i in range(n)])
    assert through.sum() > 10
    # reference walls: transmission coefficient 0 on the theta and phi components (the radial
    # component that the reference's transport leaves is passed on); slabs: the field goes through
    d = np.linalg.norm(fb - fa, axis=1)
    assert np.all(d[through] > 1e-6 * np.linalg.norm(fa, axis=1).max())
    direct = pa["pnseg"] == 1
    assert direct.any() and np.array_equal(fa[direct], fb[direct])   # no wall on the direct path
    with pytest.raises(Exception):
        gpu_ctx.trace(make_params(**dict(KW, wallWidth=-1.0)), rx)
