"""GPU parity on the cases the oracle is pinned on by hand, on the reference's own inputs, and on the
headline workload at its full size.

* The hand-worked two-bounce dihedral and the three analytic patterns of tests/test_oracle_pins.py,
  traced on the GPU and compared with the oracle (paths, corrected vertices, fields).
* The box room at its stated level 5 (10,242 launches, R 4).
* The reference's scenes with the receiver files it ships (data/RX_*.dat) and its teapot.stl
  (37,898 triangles; the BVH sort permutes the mesh, so Q1's indexing is exercised at scale).
* The urban canyon at full size (level 9, 2,621,442 launches, R 6, T 1, all 10,000 receivers)
  against the oracle's committed result on every 5th receiver, for both reception models
  (tests/golden/make_urban_full_fixture.py).  Each receiver's paths and field depend on that
  receiver only, so the full trace restricted to those receivers must equal the fixture.
"""
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params, scenes
from tests import test_oracle_pins as P
from tests.parity import FIELD_ATOL_REL, FIELD_RTOL, compare, path_table

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _both(gpu_ctx, soup, mats, kw, rx, label):
    gpu_ctx.scene_from_soup(*soup)
    gpu_ctx.materials(mats)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(O.Scene.from_soup(*soup), O.make_params(**kw), rx, mats, nthreads=8)
    worst = compare(o, g, label)
    say(label, "paths", o.np, "worst field rel err", worst)
    return o, g


@pytest.mark.parametrize("antenna", [0, 1, 2])
def test_analytic_patterns_match_oracle(gpu_ctx, antenna):
    say("direct patterns", antenna)
    s = P.Soup()
    s.tri((30, 30, 30), (31, 30, 30), (30, 31, 30), (0, 0, 1))
    rng = np.random.default_rng(9)
    d = rng.normal(size=(300, 3))
    rx = np.array([0.5, -0.3, 0.2]) + d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(5, 50, (300, 1))
    kw = dict(tx=(0.5, -0.3, 0.2), tess=4, maxR=1, maxT=0, frequency=2.4e9, pt=2.0, txCoordSys=P.TXCS,
              antennaType=antenna)
    _both(gpu_ctx, s.arrays(), np.array([[2.4e9, -1, -1, 5.0, 0.0, 0.01]]), kw, rx, f"patterns {antenna}")


def test_dihedral_two_bounce_matches_oracle(gpu_ctx):
    """All launches of level 5 in the tilted dihedral, receivers around the hand-worked launch's
    rays (inside and just outside their reception radii) and on a grid, near-perfect conductor and
    lossy walls."""
    say("dihedral")
    soup, _ = P.dihedral()
    tx = P.OFFSET + P.ROT @ P.TX_LOCAL
    rng = np.random.default_rng(4)
    grid = [P.OFFSET + P.ROT @ np.array([x, y, z]) for x in (2, 8, 20) for y in (-10, 0, 10) for z in (2, 8, 20)]
    rx = np.concatenate([np.array(grid), tx + rng.normal(size=(200, 3)) * 8.0])
    for mats in (np.array([[P.FREQ, -1, -1, 1.0, 0.0, 1e12]]), np.array([[P.FREQ, -1, -1, 4.5, 0.0, 0.05]])):
        kw = dict(tx=tx, tess=P.TESS, maxR=2, maxT=1, frequency=P.FREQ, antennaType=2, txCoordSys=P.TXCS)
        _both(gpu_ctx, soup, mats, kw, rx, "dihedral R2 T1")


def test_box_room_at_stated_level_matches_oracle(gpu_ctx):
    """The box room as the benchmark configuration states it: level 5, R 4, 100 receivers, both
    reception models."""
    cfg = scenes.config1()
    for model in (0, 1):
        say("box room L5 model", model)
        kw = dict(cfg["trace"], receptionModel=model)
        assert kw["tess"] == 5 and kw["maxR"] == 4
        _both(gpu_ctx, cfg["soup"], cfg["materials"], kw, cfg["rx"], f"box room L5 model {model}")


def read_rx(name):
    """A receiver file as the reference ships it: a title line, the count, then x y z per line
    (Preprocessor::ReadLocationFile, Preprocessor.cpp:37-67)."""
    with open(os.path.join(GOLDEN, "rx", name)) as fh:
        lines = [ln.split() for ln in fh.read().splitlines()[1:] if ln.strip()]
    n = int(lines[0][0])
    rx = np.array([[float(t) for t in ln[:3]] for ln in lines[1:1 + n]])
    assert rx.shape == (n, 3)
    return rx


RX_CASES = [
    # scene, receiver file, transmitter (the reference's main() names the office and corner ones)
    ("office.stl", "RX_Office.dat", (0.835938, 4.53906, 2.5), 3, 2, 1),
    ("corner.stl", "RX_Corner.dat", (0.0, 0.0, 2.5), 4, 3, 0),
    ("two_planes.stl", "RX_TwoPlanes.dat", (0.0, 0.0, 1.0), 4, 3, 0),
    ("ground.stl", "RX_Ground.dat", (0.0, 0.0, 5.0), 4, 1, 0),
]


@pytest.mark.parametrize("name,rxf,tx,tess,maxR,maxT", RX_CASES, ids=[c[1] for c in RX_CASES])
def test_reference_receiver_files_match_oracle(gpu_ctx, name, rxf, tx, tess, maxR, maxT):
    say("receiver file", rxf)
    rx = read_rx(rxf)
    path = os.path.join(GOLDEN, "stl", name)
    mats = np.array([[1.8e9, 221.13, 0.05, 1.0, 0.0, 1e6], [1.8e9, -1, -1, 2.5, 0.3, 0.03]])
    kw = dict(tx=tx, tess=tess, maxR=maxR, maxT=maxT, frequency=1.8e9, antennaType=1,
              txCoordSys=[[0, 1, 0], [0, 0, 1], [1, 0, 0]])
    gpu_ctx.scene_from_stl(path)
    gpu_ctx.materials(mats)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(O.Scene.from_stl(path), O.make_params(**kw), rx, mats, nthreads=8)
    compare(o, g, rxf)


def test_teapot_matches_oracle(gpu_ctx, tmp_path):
    """The reference's 37,898-triangle teapot: centre ties make the BVH sort permute the mesh, so
    the reference indexing of surfaces through the sorted vector (Q1) is exercised at scale."""
    say("teapot")
    stl = str(tmp_path / "teapot.stl")
    with gzip.open(os.path.join(GOLDEN, "stl", "teapot.stl.gz")) as fi, open(stl, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    sc = O.Scene.from_stl(stl)
    assert not np.array_equal(sc.triangles()["mesh_tri_id"], np.arange(sc.ntri))
    gpu_ctx.scene_from_stl(stl)
    mats = np.array([[2.4e9, -1, -1, 3.0, 0.0, 0.01]])
    gpu_ctx.materials(mats)
    rng = np.random.default_rng(2)
    rx = np.column_stack([rng.uniform(-4, 4, 400), rng.uniform(-4, 4, 400), rng.uniform(-0.5, 3, 400)])
    for q1 in (1, 0):
        kw = dict(tx=(2.5, -1.0, 2.0), tess=5, maxR=3, maxT=1, maxTotal=3, frequency=2.4e9, antennaType=2,
                  q1MeshIndexing=q1)
        g = gpu_ctx.trace(make_params(**kw), rx)
        o = O.trace(sc, O.make_params(**kw), rx, mats, nthreads=8)
        worst = compare(o, g, f"teapot q1 {q1}")
        say("teapot q1", q1, "paths", o.np, "bounces", o.bounces, "worst", worst)


def _fixture_compare(fx, g, label):
    """GPU result of the full receiver set restricted to the fixture's receivers, against it."""
    sel = fx["rx_index"]
    gp = g.paths()
    pos = np.full(int(sel.max()) + 1, -1)
    pos[sel] = np.arange(len(sel))
    keep = np.flatnonzero(np.isin(gp["prx"], sel))
    sub = dict(prx=pos[gp["prx"][keep]], plaunch=gp["plaunch"][keep], pnseg=gp["pnseg"][keep])
    segs = np.concatenate([np.arange(gp["poff"][k], gp["poff"][k] + gp["pnseg"][k]) for k in keep]) \
        if len(keep) else np.zeros(0, np.int64)
    sub["poff"] = np.concatenate([[0], np.cumsum(sub["pnseg"])[:-1]]) if len(keep) else np.zeros(0, np.int64)
    for k in ("stri", "scop", "stype", "ssrc", "stgt"):
        sub[k] = gp[k][segs]
    want = {k: fx[k] for k in ("prx", "plaunch", "pnseg", "poff", "stri", "scop", "stype", "ssrc", "stgt")}
    tg, to = path_table(sub), path_table(want)
    assert tg == to, f"{label}: path sets differ: {len(tg)} gpu vs {len(to)} oracle"
    assert np.array_equal(sub["ssrc"], want["ssrc"]) and np.array_equal(sub["stgt"], want["stgt"]), label
    for eg, eo, what in ((g.path_fields()[keep], fx["path_fields"], "per path"), (g.fields[sel], fx["fields"], "per rx")):
        mag = np.linalg.norm(eo, axis=1)
        err = np.linalg.norm(eg - eo, axis=1)
        atol = FIELD_ATOL_REL * (mag.max() if mag.size else 0.0)
        assert np.all(err <= FIELD_RTOL * mag + atol), f"{label} {what}: worst {np.max(err / np.maximum(mag, 1e-300))}"
    return len(to)


@pytest.mark.parametrize("model", [1, 0])
def test_full_size_urban_matches_committed_oracle_result(gpu_ctx, model):
    say("urban level 9 full size, model", model)
    fx = np.load(os.path.join(GOLDEN, f"urban_L9_full_model{model}.npz"))
    cfg = scenes.config2(tess=9)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    g = gpu_ctx.trace(make_params(**dict(cfg["trace"], receptionModel=model)), cfg["rx"])
    assert g.bounces == int(fx["bounces"])
    n = _fixture_compare(fx, g, f"urban L9 model {model}")
    say("urban L9 model", model, "fixture paths", n, "gpu paths (all receivers)", len(g.paths()["prx"]))
