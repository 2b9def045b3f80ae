"""Scene ingestion on the device (sbr_scene_from_soup_device): the triangle set equals the host
preprocessor's (vertices, file normals, coplanar ids, materials) on the reference's STL scenes and
on the synthetic configurations up to the 1M-triangle city; the linear BVH is well formed (every
triangle in one leaf, every box the exact bound of what it holds); and traces over it equal the
oracle's reference traversal over the same exported tree (oracle Scene.from_arrays)."""
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, make_params, scenes
from tests.parity import compare

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _load_soup(name, tmp_path):
    if name == "teapot.stl":
        p = str(tmp_path / name)
        with gzip.open(os.path.join(GOLDEN, "stl", "teapot.stl.gz")) as fi, open(p, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        return p
    return os.path.join(GOLDEN, "stl", name)


def check_bvh(ctx):
    t = ctx.scene_triangles()
    b = ctx.scene_bvh_dfs()
    n = len(t["coplanar"])
    ids = b["tri_dfs_id"]
    assert np.array_equal(np.sort(ids), np.arange(n))
    v = t["verts"].reshape(-1, 3, 3)
    tlo, thi = v.min(1), v.max(1)
    leaf = b["leaf"] == 1
    covered = 0
    for k in np.flatnonzero(leaf):
        sel = ids[b["a"][k]:b["a"][k] + b["b"][k]]
        covered += len(sel)
        assert np.array_equal(b["box"][k, :3], tlo[sel].min(0)) and np.array_equal(b["box"][k, 3:], thi[sel].max(0))
    assert covered == n
    for k in np.flatnonzero(~leaf):
        c0, c1 = b["a"][k], b["b"][k]
        assert c0 == k + 1 and c1 > c0
        assert np.array_equal(b["box"][k, :3], np.minimum(b["box"][c0, :3], b["box"][c1, :3]))
        assert np.array_equal(b["box"][k, 3:], np.maximum(b["box"][c0, 3:], b["box"][c1, 3:]))
    return t, b


@pytest.mark.parametrize("name", ["office.stl", "bahen.stl", "menger_sponge.stl", "two_icospheres.stl", "teapot.stl"])
def test_device_ingest_of_reference_scenes(gpu_ctx, name, tmp_path):
    say("ingest", name)
    path = _load_soup(name, tmp_path)
    with Context(-1) as host:
        host.scene_from_stl(path)
        want = host.scene_triangles()
    # the facets of the file: the host reader's corner order is the file's, so re-read them
    import struct
    raw = open(path, "rb").read()
    if raw[:5].lower() == b"solid" and b"facet" in raw[:400]:
        nrm, cor = [], []
        for ln in raw.decode().splitlines():
            w = ln.split()
            if w[:2] == ["facet", "normal"]:
                nrm.append([float(x) for x in w[2:5]])
            elif w[:1] == ["vertex"]:
                cor.append([float(x) for x in w[1:4]])
        c9, n3 = np.array(cor, np.float32).reshape(-1, 9), np.array(nrm, np.float32)
    else:
        cnt = struct.unpack("<I", raw[80:84])[0]
        rec = np.frombuffer(raw[84:84 + 50 * cnt], dtype=np.dtype([("f", "<f4", 12), ("a", "<u2")]))
        n3, c9 = rec["f"][:, :3].copy(), rec["f"][:, 3:].copy()
    gpu_ctx.scene_from_soup_device(c9, n3)
    got = gpu_ctx.scene_triangles()
    for k in ("verts", "normals", "coplanar", "material"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["mesh_tri_id"], np.arange(len(got["coplanar"])))
    check_bvh(gpu_ctx)


def _oracle_on_device_tree(ctx):
    t = ctx.scene_triangles()
    b = ctx.scene_bvh_dfs()
    return O.Scene.from_arrays(t, b)


@pytest.mark.parametrize("model", [0, 1])
def test_trace_on_device_tree_matches_oracle(gpu_ctx, model):
    cfg = scenes.config2(tess=4)
    say("urban on the device-built tree, model", model)
    gpu_ctx.scene_from_soup_device(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    check_bvh(gpu_ctx)
    sc = _oracle_on_device_tree(gpu_ctx)
    kw = dict(cfg["trace"], receptionModel=model)
    g = gpu_ctx.trace(make_params(**kw), cfg["rx"])
    o = O.trace(sc, O.make_params(**kw), cfg["rx"], cfg["materials"], nthreads=16)
    worst = compare(o, g, f"device tree model {model}")
    say("paths", o.np, "bounces", o.bounces, "worst", worst)
    # the device tree changes no discrete result of this scene's trace against the SAH tree's
    with Context(0) as c2:
        c2.scene_from_soup(*cfg["soup"])
        c2.materials(cfg["materials"])
        s = c2.trace(make_params(**kw), cfg["rx"])
        assert s.bounces == g.bounces


def test_city_ingest_and_trace(gpu_ctx):
    name, soup, boxes, half, mats = scenes.config5_scene()
    ms = gpu_ctx.scene_from_soup_device(*soup)
    say(name, "device ingest ms", ms)
    gpu_ctx.materials(mats)
    with Context(-1) as host:
        host.scene_from_soup(*soup)
        want = host.scene_triangles()
    got = gpu_ctx.scene_triangles()
    for k in ("verts", "normals", "coplanar", "material"):
        assert np.array_equal(got[k], want[k]), k
    check_bvh(gpu_ctx)
    sc = _oracle_on_device_tree(gpu_ctx)
    rng = np.random.default_rng(8)
    for kw, _ in scenes.coverage_queries(3, boxes, seed=20261203 + 77, half=0.5 * half, tess=(5, 5), max_rx=1):
        c = np.asarray(kw["tx"])
        rx = np.column_stack([c[0] + rng.uniform(-150, 150, 150), c[1] + rng.uniform(-150, 150, 150), np.full(150, 1.5)])
        g = gpu_ctx.trace(make_params(**kw), rx)
        o = O.trace(sc, O.make_params(**kw), rx, mats, nthreads=16)
        compare(o, g, "city on the device tree")
        say("city query paths", o.np, "bounces", o.bounces)
