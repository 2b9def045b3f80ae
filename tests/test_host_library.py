"""CPU checks of the product library without a GPU: it loads, exports every symbol the headers
declare, and its host scene preparation (STL parse, vertex dedup, coplanar ids, SAH BVH and the
mesh permutation) equals the oracle's."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import Context, SbrError, lib, make_params, scenes
from sbrsolver_amd.capi import SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sbr_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = lib()
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert decl == set(SIGNATURES), decl ^ set(SIGNATURES)


def test_host_only_context_rejects_device_work():
    with Context(-1) as c:
        c.scene_from_stl(os.path.join(GOLDEN, "stl", "office.stl"))
        c.materials()
        with pytest.raises(SbrError):
            c.trace(make_params(tx=(0, 0, 1), tess=1, maxR=1), np.zeros((1, 3)))
        with pytest.raises(SbrError, match="no device"):
            c.autotune(make_params(tx=(0, 0, 1), tess=1, maxR=1), np.zeros((1, 3)))
        with pytest.raises(SbrError, match="out of range"):
            c.set_tuning(window_nodes=-5)
        c.set_tuning(1, 2, 1 << 20, 6, 3, 5000)
        assert c.get_tuning() == dict(lds_scene=1, blocks_per_cu=2, window_nodes=1 << 20, stack_lds=6, oversub=3,
                                      table_entries=5000)
        for bad in (dict(stack_lds=-1), dict(stack_lds=65), dict(oversub=-2), dict(oversub=65), dict(table_entries=-1),
                    dict(table_entries=(1 << 32) + 1)):
            with pytest.raises(SbrError):
                c.set_tuning(**bad)
        with pytest.raises(SbrError):
            c.scene_from_stl("/nonexistent.stl")


def test_failed_upload_leaves_no_scene():
    """A scene upload that fails leaves the context without a scene (never a mix of two scenes):
    later calls that need one report SBR_E_STATE."""
    import ctypes as C
    with Context(-1) as c:
        c.scene_from_stl(os.path.join(GOLDEN, "stl", "office.stl"))
        assert c.scene_info()["n_tri"] > 0
        desc = (C.c_char * 512)()   # zeroed: no triangles, no arrays
        assert lib().sbr_scene_upload(c.h, desc) == -1   # SBR_E_ARG
        with pytest.raises(SbrError, match="no scene"):
            c.scene_info()
        c.scene_from_stl(os.path.join(GOLDEN, "stl", "cube.stl"))   # a good upload restores it
        assert c.scene_info()["n_tri"] == 12


def _dfs_leaf_order(ob):
    """DFS (children[0] first) leaf order of the oracle's BVH = the device triangle order."""
    return ob["leaf_tri"]


SCENES = ["ground.stl", "corner.stl", "office.stl", "bahen.stl", "menger_sponge.stl", "two_icospheres.stl",
          "geometry_test2.stl"]


@pytest.mark.parametrize("name", SCENES)
def test_host_scene_matches_oracle(name):
    path = os.path.join(GOLDEN, "stl", name)
    with Context(-1) as c:
        c.scene_from_stl(path)
        mine = c.scene_triangles()
        bvh = c.scene_bvh_dfs()
    sc = O.Scene.from_stl(path)
    ref = sc.triangles()
    for k in ("coplanar", "material", "verts", "normals", "mesh_tri_id"):
        assert np.array_equal(mine[k], ref[k]), k
    ob = sc.bvh()
    assert np.array_equal(bvh["tri_dfs_id"], _dfs_leaf_order(ob))
    assert bvh["box"].shape[0] == ob["box"].shape[0]
    # same set of node boxes (creation order vs DFS order)
    a = np.array(sorted(map(tuple, bvh["box"])))
    b = np.array(sorted(map(tuple, ob["box"])))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("builder", ["c1", "c2"])
def test_host_synthetic_scenes_match_oracle(builder):
    cfg = scenes.config1() if builder == "c1" else scenes.config2(tess=2)
    with Context(-1) as c:
        c.scene_from_soup(*cfg["soup"])
        mine = c.scene_triangles()
        bvh = c.scene_bvh_dfs()
    sc = O.Scene.from_soup(*cfg["soup"])
    ref = sc.triangles()
    for k in ("coplanar", "material", "verts", "normals", "mesh_tri_id"):
        assert np.array_equal(mine[k], ref[k]), k
    assert np.array_equal(bvh["tri_dfs_id"], sc.bvh()["leaf_tri"])


def test_config2_shape():
    cfg = scenes.config2(tess=9)
    v, n, m = cfg["soup"]
    assert v.shape == (482, 9) and n.shape == (482, 3) and m.shape == (482,)
    assert cfg["rx"].shape == (10000, 3)
    assert np.allclose(np.abs(n).sum(axis=1), 1.0)


def test_oracle_over_an_exported_tree_equals_its_own_build():
    """The oracle traced over the library's exported DFS tree (Scene.from_arrays) gives its own
    result: the same tree in another node numbering, so every trace and path is identical."""
    cfg = scenes.config2(tess=3)
    with Context(-1) as c:
        c.scene_from_soup(*cfg["soup"])
        t = c.scene_triangles()
        b = c.scene_bvh_dfs()
    a = O.Scene.from_arrays(t, b)
    own = O.Scene.from_soup(*cfg["soup"])
    kw = dict(cfg["trace"], tess=3)
    ra = O.trace(a, O.make_params(**kw), cfg["rx"][::7], cfg["materials"], nthreads=8)
    ro = O.trace(own, O.make_params(**kw), cfg["rx"][::7], cfg["materials"], nthreads=8)
    assert ra.bounces == ro.bounces and ra.np == ro.np > 0
    assert np.array_equal(ra.stri, ro.stri) and np.array_equal(ra.ssrc, ro.ssrc)
    assert np.array_equal(ra.fields, ro.fields, equal_nan=True)
