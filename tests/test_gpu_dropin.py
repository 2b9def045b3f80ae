"""Drop-in check: the reference's main() (SBRSolver_rev2.cpp:18-144) rebuilt against the library
(sbrsolver_amd/sbr_demo, C++ RTSolver / Preprocessor / BVH call surface) writes the same files the
reference run committed under data/output/."""
import os
import subprocess

import numpy as np
import pytest

from tests import fixtures_io

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DEMO = os.path.join(ROOT, "sbrsolver_amd", "sbr_demo")


@pytest.fixture(scope="module")
def demo_out(tmp_path_factory):
    out = tmp_path_factory.mktemp("demo")
    print("[case] sbr_demo", flush=True)
    r = subprocess.run([DEMO, os.path.join(GOLDEN, "stl", "ground.stl"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return str(out)


def test_field_csv(demo_out):
    mine = open(os.path.join(demo_out, "ElectricField.csv")).read().split("\n")
    ref = open(os.path.join(GOLDEN, "output", "ElectricField.csv")).read().split("\n")
    assert len(mine) == len(ref)
    assert mine[0] == ref[0] and mine[1].split(",")[:3] == ref[1].split(",")[:3]   # showpos from row 2 on
    assert mine[2].split(",")[:3] == ref[2].split(",")[:3]
    xyz_m, e_m = fixtures_io.read_field_csv(os.path.join(demo_out, "ElectricField.csv"))
    xyz_r, e_r = fixtures_io.read_field_csv(os.path.join(GOLDEN, "output", "ElectricField.csv"))
    assert np.array_equal(xyz_m, xyz_r)
    rel = np.linalg.norm(e_m - e_r, axis=1) / np.linalg.norm(e_r, axis=1)
    assert rel.max() < 2e-5
    same = sum(a == b for a, b in zip(mine, ref))
    assert same >= 0.95 * len(ref), f"only {same} of {len(ref)} lines byte-identical"


def test_raypath_vtk(demo_out):
    m = fixtures_io.read_raypath_vtk(os.path.join(demo_out, "RayPath.vtk"))
    r = fixtures_io.read_raypath_vtk(os.path.join(GOLDEN, "output", "RayPath.vtk"))
    assert np.array_equal(m["launch"], r["launch"]) and np.array_equal(m["rx"], r["rx"])
    assert m["lines"] == r["lines"]
    assert np.allclose(m["points"], r["points"], rtol=1e-5, atol=1e-4)
    assert np.allclose(m["eabs"], r["eabs"], rtol=2e-5)
    assert np.allclose(m["phase"], r["phase"], atol=5e-5)


def test_rxlocations_vtk(demo_out):
    m = fixtures_io.read_rxloc_vtk(os.path.join(demo_out, "RxLocations.vtk"))
    r = fixtures_io.read_rxloc_vtk(os.path.join(GOLDEN, "output", "RxLocations.vtk"))
    assert np.allclose(m["eabs"], r["eabs"], rtol=2e-5)
    assert np.allclose(m["edb"], r["edb"], atol=2e-4)
    assert np.allclose(m["phase"], r["phase"], atol=5e-5)


def test_icosahedron_vtk_byte_identical(demo_out):
    mine = open(os.path.join(demo_out, "Icosahedron.vtk")).read()
    ref = open(os.path.join(GOLDEN, "output", "Icosahedron.vtk")).read()
    assert mine == ref
