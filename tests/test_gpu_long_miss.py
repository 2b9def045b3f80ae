"""Long misses (longMiss = 1) on the GPU: the same trees, kept paths and fields as the oracle on the
flat ground of the closed-form test and on the urban scene."""
import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params, scenes
from tests.parity import compare
from tests.test_long_miss import H, Z, ground, ground_soup

pytestmark = pytest.mark.gpu


def test_long_miss_ground_matches_oracle(gpu_ctx):
    print("[case] flat ground, long misses", flush=True)
    sc = ground()
    gpu_ctx.scene_from_soup(*ground_soup())
    mats = np.array([[1.8e9, 5.0, 0.0, 5.0, 0.0, 0.01]])
    gpu_ctx.materials(mats)
    rx = np.array([[x, 0.0, Z] for x in (40.0, 80.0, 150.0, 260.0, 400.0, 700.0)])
    kw = dict(tx=(0.0, 0.0, H), tess=7, maxR=1, maxT=0, frequency=1.8e9, receptionModel=1, longMiss=1)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, mats, nthreads=8)
    compare(o, g, "ground long misses")


def test_long_miss_urban_matches_oracle(gpu_ctx):
    print("[case] urban L3 R6T1 model 1, long misses", flush=True)
    cfg = scenes.config2(tess=3)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    tr = dict(cfg["trace"], receptionModel=1, longMiss=1)
    g = gpu_ctx.trace(make_params(**tr), cfg["rx"])
    o = O.trace(O.Scene.from_soup(*cfg["soup"]), O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    compare(o, g, "urban long misses")
