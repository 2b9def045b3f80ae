"""The reference's C++ call surface without a GPU (tests/surface_check.cpp, compiled against
include/): Preprocessor::StlToGeometry(..., saveEdges, saveFaces, dir) and BVH<Triangle>::SaveAsVtk
write the reference's Edges.vtk, Faces.vtk and BVH.vtk for ground.stl byte for byte; the path-tree
helpers of Ray.h (newPathTreeNode numbering, CloneNode / CloneTree, Paths(root), DeleteChildNodes)
behave as Ray.cpp; ReadLocationFile reads a receiver file as the reference ships it; FieldCompute
with no paths gives zero without touching a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sbrsolver_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libsbr_amd.so")):
        pytest.skip("library not built")
    d = tmp_path_factory.mktemp("surf")
    exe = str(d / "surface_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "surface_check.cpp"), "-o", exe, "-L", LIBDIR, "-lsbr_amd",
                           "-Wl,-rpath," + LIBDIR])
    out = subprocess.check_output([exe, os.path.join(GOLDEN, "stl", "ground.stl"), str(d),
                                   os.path.join(GOLDEN, "rx", "RX_Ground.dat")], text=True)
    facts = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    return d, facts


@pytest.mark.parametrize("name", ["Edges.vtk", "Faces.vtk"])
def test_vtk_writers_reproduce_reference_files(run, name):
    d, facts = run
    assert facts["stl"] == ["1", "2", "1"]
    with open(os.path.join(d, name), "rb") as a, open(os.path.join(GOLDEN, "output", name), "rb") as b:
        assert a.read() == b.read()


def test_bvh_vtk_matches_reference_layout(run):
    """The reference's BVH.vtk was written for an earlier ground.stl (a 20 x 20 m square; the
    committed one spans x -5..5, y -1..1000): every line but the 8 box corners is byte-identical,
    and the corners are the ground's box in the reference's order (x fastest, then y, then z)."""
    d, _ = run
    mine = open(os.path.join(d, "BVH.vtk")).read().split("\n")
    ref = open(os.path.join(GOLDEN, "output", "BVH.vtk")).read().split("\n")
    assert len(mine) == len(ref)
    pts = slice(6, 14)
    assert mine[:6] == ref[:6] and mine[14:] == ref[14:]
    lo, hi = (-5, -1, 0), (5, 1000, 0)
    want = [f" {(lo, hi)[k & 1][0]} {(lo, hi)[(k >> 1) & 1][1]} {(lo, hi)[(k >> 2) & 1][2]}" for k in range(8)]
    assert mine[pts] == want
    assert [ln.split() for ln in ref[pts]] == [[str((-10, 10)[k & 1]), str((-10, 10)[(k >> 1) & 1]), "0"]
                                              for k in range(8)]


def test_path_tree_helpers(run):
    _, f = run
    assert f["ids"] == ["0", "1", "2", "3", "4", "5", "next", "6"]
    assert " ".join(f["paths"]) == "2 [ -1 12 13 ] [ -1 10 11 ]"
    assert f["clone"] == ["1", "1", "1"]
    assert f["deleted"] == ["1", "1"] and f["empty"] == ["0"]


def test_receiver_file_and_empty_field(run):
    _, f = run
    assert f["rx"] == ["1", "1", "0", "40", "3"]
    assert f["fc"] == ["0", "0"]
