"""Physical field transport (physicalField = 1) on the GPU: the same paths and fields as the oracle
(whose transport is pinned by image theory in test_physical_field.py) on the flat ground, the office
scene (transmissions, slab walls) and the urban scene."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params, scenes
from tests.parity import compare
from tests.office import KW, MATS, STL
from tests.office import scene_and_rx as _rx
from tests.test_long_miss import ground_soup
from tests.test_physical_field import FREQ, H, HORZ

pytestmark = pytest.mark.gpu


def test_physical_field_ground_matches_oracle(gpu_ctx):
    print("[case] ground, horizontal dipole, physical field", flush=True)
    gpu_ctx.scene_from_soup(*ground_soup())
    mats = np.array([[FREQ, 15.0, 0.0, 15.0, 0.0, 0.5]])
    gpu_ctx.materials(mats)
    rx = np.array([[0.0, 60.0, 2.0], [90.0, -70.0, 1.5], [300.0, 200.0, 20.0]])
    kw = dict(tx=(0.0, 0.0, H), tess=8, maxR=1, maxT=0, frequency=FREQ, receptionModel=1, longMiss=1,
              txCoordSys=HORZ, physicalField=1)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(O.Scene.from_soup(*ground_soup()), O.make_params(**kw), rx, mats, nthreads=8)
    compare(o, g, "ground physical field")


def test_physical_field_office_matches_oracle(gpu_ctx):
    print("[case] office L4 R2T2, physical field, slab walls", flush=True)
    sc, rx = _rx()
    gpu_ctx.scene_from_stl(STL)
    gpu_ctx.materials(MATS)
    kw = dict(KW, tess=4, maxT=2, q1MeshIndexing=0, wallWidth=0.12, physicalField=1)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, MATS, nthreads=8)
    compare(o, g, "office physical field")


def test_physical_field_urban_matches_oracle(gpu_ctx):
    print("[case] urban L3 R6T1 model 1, physical field", flush=True)
    cfg = scenes.config2(tess=3)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    tr = dict(cfg["trace"], receptionModel=1, physicalField=1)
    g = gpu_ctx.trace(make_params(**tr), cfg["rx"])
    o = O.trace(O.Scene.from_soup(*cfg["soup"]), O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    compare(o, g, "urban physical field")


def test_physical_mode_urban_matches_oracle(gpu_ctx):
    print("[case] urban L3 R6T1, every physical switch, 0.2 m slab walls", flush=True)
    cfg = scenes.config2(tess=3)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    p = make_params(**dict(cfg["trace"], physical=True, wallWidth=0.2))
    assert (p.q1MeshIndexing, p.receptionModel, p.typedDedup, p.checkOcclusion, p.longMiss, p.physicalField) == \
        (0, 1, 1, 1, 1, 1)
    g = gpu_ctx.trace(p, cfg["rx"])
    tr = dict(cfg["trace"], q1MeshIndexing=0, receptionModel=1, typedDedup=1, checkOcclusion=1, longMiss=1,
              physicalField=1, wallWidth=0.2)
    o = O.trace(O.Scene.from_soup(*cfg["soup"]), O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    compare(o, g, "urban physical mode")
