"""Parity of the HIP path (through the C-ABI) with the FP64 oracle, on the same inputs.

Discrete results are compared exactly: launch directions (bit-exact), closest-hit triangle and
point, interaction trees node by node, kept paths per receiver as ordered triangle-id sequences
with bit-identical corrected vertices, bounce counts.  Fields within 1e-4 relative per receiver.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import make_params, scenes
from tests import fixtures_io
from tests.parity import compare

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("tess", [0, 1, 2, 3, 5, 7])
def test_raygen_bit_exact(gpu_ctx, tess):
    say("raygen", tess)
    g = gpu_ctx.generate_rays(tess)
    o = O.icosahedron(tess)
    assert g.shape == o.shape == (10 * 4 ** tess + 2, 3)
    assert np.array_equal(g, o)


def test_raygen_level9_bit_exact(gpu_ctx):
    say("raygen 9")
    g = gpu_ctx.generate_rays(9)
    o = O.icosahedron(9)
    assert np.array_equal(g, o)


def _stl(name):
    return os.path.join(GOLDEN, "stl", name)


def _rand_rays(n, seed, center, spread):
    rng = np.random.default_rng(seed)
    org = center + spread * (rng.random((n, 3)) - 0.5)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org, d


@pytest.mark.parametrize("name", ["ground.stl", "cube.stl", "office.stl", "bahen.stl", "menger_sponge.stl",
                                  "two_icospheres.stl"])
@pytest.mark.parametrize("exact", [0, 1, 2])
def test_closest_hit_matches_oracle(gpu_ctx, name, exact):
    say("closest_hit", name, exact)
    gpu_ctx.scene_from_stl(_stl(name))
    sc = O.Scene.from_stl(_stl(name))
    tr = sc.triangles()
    vmin, vmax = tr["verts"].reshape(-1, 3).min(0), tr["verts"].reshape(-1, 3).max(0)
    org, d = _rand_rays(2000, 7, (vmin + vmax) / 2, (vmax - vmin) * 1.2 + 1e-3)
    tri, pt, dist = gpu_ctx.closest_hit(org, d, exact=exact)
    for i in range(len(d)):
        ot, op, od = sc.closest_hit(org[i], d[i])
        assert tri[i] == ot, (i, tri[i], ot)
        if ot >= 0:
            assert np.array_equal(pt[i], op) and dist[i] == od


def _trees_equal(gres, otrees, label):
    g = gres.trees()
    for k in ("launch", "depth", "code", "tri", "cop", "type"):
        assert np.array_equal(g[k], otrees[k]), f"{label}: tree field {k} differs"
    for k in ("src", "tgt", "pl"):
        assert np.array_equal(g[k], otrees[k]), f"{label}: tree field {k} differs"


def test_ground_fixture_reproduced(gpu_ctx):
    """The committed main() run: GPU fields against data/output/ElectricField.csv (6 digits)."""
    say("ground fixture")
    cfg = scenes.ground_fixture()
    gpu_ctx.scene_from_stl(_stl("ground.stl"))
    gpu_ctx.materials(cfg["materials"])
    res = gpu_ctx.trace(make_params(**cfg["trace"]), cfg["rx"])
    xyz, E = fixtures_io.read_field_csv(os.path.join(GOLDEN, "output", "ElectricField.csv"))
    rel = np.linalg.norm(res.fields - E, axis=1) / np.linalg.norm(E, axis=1)
    assert rel.max() < 2e-5
    p = res.paths()
    assert len(p["prx"]) == 3133
    vtk = fixtures_io.read_raypath_vtk(os.path.join(GOLDEN, "output", "RayPath.vtk"))
    order = np.lexsort((p["prx"], p["plaunch"]))   # the VTK lists paths by (launch, receiver)
    assert np.array_equal(p["plaunch"][order], vtk["launch"])
    assert np.array_equal(p["prx"][order], vtk["rx"])


CASES = [
    # (label, scene builder, trace overrides, receivers)
    ("box_room_L3_R4_model0", "c1", dict(tess=3), None),
    ("box_room_L3_R4_model1", "c1", dict(tess=3, receptionModel=1), None),
    ("urban_L3_R6T1_model0", "c2", dict(tess=3), None),
    ("urban_L4_R6T1_model1", "c2", dict(tess=4, receptionModel=1), None),
    ("urban_L3_exact_traversal", "c2", dict(tess=3, exactTraversal=1), None),
    ("urban_L3_reference_tree_clipped", "c2", dict(tess=3, exactTraversal=2), None),
    ("urban_L3_physical_indexing", "c2", dict(tess=3, q1MeshIndexing=0), None),
    ("urban_L3_total_cap3", "c2", dict(tess=3, maxTotal=3), None),
    ("urban_L3_isotropic_nofresnel", "c2", dict(tess=3, antennaType=0, useFresnel=False), None),
    ("urban_L3_slab_walls_model1", "c2", dict(tess=3, receptionModel=1, wallWidth=0.2), None),
]


def _scene(kind):
    return scenes.config1() if kind == "c1" else scenes.config2(tess=3)


@pytest.mark.parametrize("label,kind,over,_", CASES, ids=[c[0] for c in CASES])
def test_trace_matches_oracle(gpu_ctx, label, kind, over, _):
    say("trace", label)
    cfg = _scene(kind)
    tr = dict(cfg["trace"])
    tr.update(over)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    g = gpu_ctx.trace(make_params(**tr), cfg["rx"])
    sc = O.Scene.from_soup(*cfg["soup"])
    otr = {k: v for k, v in tr.items() if k != "exactTraversal"}
    o = O.trace(sc, O.make_params(**otr), cfg["rx"], cfg["materials"], nthreads=8)
    _trees_equal(g, O.trees(sc, O.make_params(**otr)), label)
    worst = compare(o, g, label)
    say(label, "paths", len(o.prx), "bounces", o.bounces, "worst field rel err", worst)


STL_CASES = [
    ("corner.stl", (0.0, 0.0, 2.5), 3, 2, 1),
    ("two_planes.stl", (0.0, 0.0, 1.0), 3, 3, 0),
    ("office.stl", (0.835938, 4.53906, 2.5), 2, 2, 1),
    ("geometry_test2.stl", (0.1, 0.2, 0.5), 3, 2, 1),
]


@pytest.mark.parametrize("name,tx,tess,maxR,maxT", STL_CASES, ids=[c[0] for c in STL_CASES])
def test_stl_scenes_match_oracle(gpu_ctx, name, tx, tess, maxR, maxT):
    say("stl trace", name)
    gpu_ctx.scene_from_stl(_stl(name))
    sc = O.Scene.from_stl(_stl(name))
    tr = sc.triangles()
    v = tr["verts"].reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    g1 = np.linspace(lo[0], hi[0], 9)
    g2 = np.linspace(lo[1], hi[1], 9)
    rx = np.array([[a, b, (lo[2] + hi[2]) / 2 + 0.1] for a in g1 for b in g2])
    mats = np.array([[1.8e9, -1, -1, 2.5, 0.3, 0.03], [1.8e9, -1, -1, 6.0, 0.05, 0.005]])
    gpu_ctx.materials(mats)
    kw = dict(tx=tx, tess=tess, maxR=maxR, maxT=maxT, frequency=1.8e9, antennaType=2)
    g = gpu_ctx.trace(make_params(**kw), rx)
    o = O.trace(sc, O.make_params(**kw), rx, mats, nthreads=8)
    _trees_equal(g, O.trees(sc, O.make_params(**kw)), name)
    compare(o, g, name)


def test_batch_equals_individual_traces(gpu_ctx):
    """Transmitters traced together give, each, exactly what a trace of that transmitter alone gives."""
    say("batch vs individual")
    cfg = scenes.config2(tess=3)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    base = dict(cfg["trace"])
    variants = [dict(base, tess=3), dict(base, tess=4, maxR=4, receptionModel=1),
                dict(base, tess=2, tx=(10.0, -60.0, 12.0), maxT=0), dict(base, tess=3, launchStride=3, launchOffset=1)]
    params = [make_params(**v) for v in variants]
    batch = gpu_ctx.trace_batch(params, cfg["rx"])
    bp = batch.paths()
    btx = batch.path_tx()
    bpf = batch.path_fields()
    bf = batch.fields   # results are read before the next trace on the context
    total = 0
    for t, p in enumerate(params):
        single = gpu_ctx.trace(p, cfg["rx"])
        total += single.bounces
        sp = single.paths()
        sel = np.flatnonzero(btx == t)
        assert np.array_equal(bp["prx"][sel], sp["prx"]) and np.array_equal(bp["plaunch"][sel], sp["plaunch"])
        assert np.array_equal(bp["pnseg"][sel], sp["pnseg"])
        assert np.array_equal(bf[t], single.fields)
        assert np.array_equal(bpf[sel], single.path_fields())
    assert batch.bounces == total


def test_full_level_urban_sample_matches_oracle(gpu_ctx):
    """The headline workload at its full size (urban canyon, icosahedron level 9, R6 T1, 10k
    receivers): every launch traced on the GPU gives the oracle's bounce count (committed fixture),
    and every 512th launch (5120 of the 2.6 M level-9 directions) gives the oracle's trees, kept
    paths and fields, with the reference's reception sphere (thousands of paths) and with
    alpha*L/sqrt(3) (a handful: its radius is that of the full level)."""
    import json
    say("urban level 9: full bounce count, sampled parity")
    cfg = scenes.config2(tess=9)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    with open(os.path.join(GOLDEN, "traversal_counts.json")) as fh:
        fixture = json.load(fh)["urban_canyon_482tri_L9_R6T1"]
    full = gpu_ctx.trace(make_params(**dict(cfg["trace"], receptionModel=1)), cfg["rx"])
    assert full.bounces == fixture["bounces"]
    sc = O.Scene.from_soup(*cfg["soup"])
    for model in (0, 1):
        tr = dict(cfg["trace"], receptionModel=model, launchStride=512, launchOffset=7)
        g = gpu_ctx.trace(make_params(**tr), cfg["rx"])
        if model == 0:
            _trees_equal(g, O.trees(sc, O.make_params(**tr)), "urban L9 sample")
        o = O.trace(sc, O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=16)
        worst = compare(o, g, f"urban L9 sample model {model}")
        say("urban L9 sample model", model, "paths", len(o.prx), "bounces", o.bounces, "worst field rel err", worst)
