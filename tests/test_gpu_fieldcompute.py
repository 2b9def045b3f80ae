"""FieldCompute and the field entry points on the GPU.

* tests/fieldcompute_check.cpp (compiled against include/): FieldCompute built over the paths an
  RTSolver kept gives RTSolver's fields bit for bit, per receiver and all at once; the kept paths'
  Ray fields; RefCoeffTest / TransCoeffTest files; a mesh and BVH edited in place are traced anew.
* sbr_path_fields on the trace's own kept paths equals the trace's per-path fields bit for bit.
* sbr_fresnel_sweep (RefCoeffTest's sweep) against the oracle's coefficients with air on the
  incident side, and against closed forms between two dielectrics: normal incidence
  (n_t - n_i) / (n_t + n_i) in the code's sign convention, the Brewster angle atan(n_t / n_i), total
  reflection beyond the critical angle.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import lib, make_params, scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sbrsolver_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cpp(tmp_path_factory):
    d = tmp_path_factory.mktemp("fc")
    exe = str(d / "fieldcompute_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "fieldcompute_check.cpp"), "-o", exe, "-L", LIBDIR,
                           "-lsbr_amd", "-Wl,-rpath," + LIBDIR])
    say("fieldcompute_check")
    r = subprocess.run([exe, os.path.join(GOLDEN, "stl", "office.stl"), str(d)], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    facts = {ln.split()[0]: [int(x) for x in ln.split()[1:]] for ln in r.stdout.splitlines() if ln.strip()}
    return d, facts


def test_fieldcompute_equals_solver_fields(cpp):
    _, f = cpp
    n, nrx, filled = f["trace"]
    assert nrx >= 250 and n == 642 * nrx and filled == 1
    equal, equal_all, nonzero, status = f["fieldcompute"]
    assert status == 0 and equal == nrx and equal_all == nrx and nonzero > nrx // 2


def test_kept_paths_ray_fields(cpp):
    _, f = cpp
    paths, multi, cap, length, angle, ids, types = f["rays"]
    assert paths > 100 and multi > 50
    assert cap == paths and length == paths and angle == paths and ids == paths
    assert types == 3   # reflections and transmissions both present


def test_mesh_edited_in_place_is_traced_anew(cpp):
    _, f = cpp
    n2, changed, n3, restored = f["inplace"]
    nrx = f["trace"][1]
    assert n2 > 0 and n3 > 0 and changed > nrx // 2 and restored == nrx


def test_coefficient_test_files(cpp):
    d, f = cpp
    assert f["coeff"] == [0, 0]
    ref = open(os.path.join(d, "ref_coeff.csv")).read().splitlines()
    tra = open(os.path.join(d, "trans_coeff.csv")).read().splitlines()
    assert ref[0] == "theta_i,refTM_mag,refTM_phase,refTE_mag, refTE_phase" and len(ref) == 92
    assert tra[0] == "theta_i,tranTM_mag,tranTM_phase, tranTE_mag, tranTE_phase" and len(tra) == 92
    rows = np.array([[float(x) for x in ln.split(",")] for ln in ref[1:]])
    assert rows[0, 0] == 0 and rows[-1, 0] == 90 and np.allclose(np.diff(rows[:, 0]), 1.0)
    eps0 = 1.0 / (4 * np.pi * 1e-7 * 2.99792458e8 ** 2)
    er, sig, fr = 5.31, 0.0326, 2.4e9
    for i in (0, 17, 45, 80):
        c = O.fresnel(np.deg2rad(rows[i, 0]), er, sig, fr)
        # the checker's conductivity term uses eps0 = 8.854187817e-12: relative differences ~1e-10
        assert abs(rows[i, 1] - abs(c["rTM"])) < 1e-5 and abs(rows[i, 3] - abs(c["rTE"])) < 1e-5
    assert eps0 > 0
    trows = np.array([[float(x) for x in ln.split(",")] for ln in tra[1:]])
    assert np.all(trows[:, 1] == 0) and np.all(trows[:, 3] == 0)   # infinite wall: no transmission


def sweep(gpu_ctx, n, freq, eps_i, eps_t, width=0.0, inf_wall=1):
    th = np.zeros(n)
    out = np.zeros((n, 4, 2))
    ei = np.array([eps_i.real, eps_i.imag])
    et = np.array([eps_t.real, eps_t.imag])
    rc = lib().sbr_fresnel_sweep(gpu_ctx.h, n, freq, ei.ctypes.data, et.ctypes.data, width, inf_wall,
                                 th.ctypes.data, out.ctypes.data)
    assert rc == 0, lib().sbr_last_error(gpu_ctx.h)
    c = out[..., 0] + 1j * out[..., 1]
    return th, dict(rTM=c[:, 0], tTM=c[:, 1], rTE=c[:, 2], tTE=c[:, 3])


def test_fresnel_sweep_against_oracle_and_closed_forms(gpu_ctx):
    say("fresnel sweep")
    eps0 = 1.0 / (4 * np.pi * 1e-7 * 2.99792458e8 ** 2)
    f = 2.4e9
    for er, sig in ((5.31, 0.0326), (2.5, 0.03), (1.0, 1e6)):
        th, c = sweep(gpu_ctx, 181, f, 1 + 0j, complex(er, -sig / (eps0 * 2 * np.pi * f)))
        assert th[0] == 0 and th[-1] == np.pi / 2
        for i in range(0, 181, 9):
            o = O.fresnel(th[i], er, sig, f)
            for k in ("rTM", "rTE"):
                assert abs(c[k][i] - o[k]) <= 1e-12 * max(1.0, abs(o[k])), (er, sig, i, k)
        assert np.all(c["tTM"] == 0) and np.all(c["tTE"] == 0)
        o = O.fresnel(0.4, er, sig, f, width=0.0, inf_wall=False)   # the slab branch at width 0
    # two dielectrics: normal incidence, Brewster, total reflection
    ni, nt = 1.5, 2.2
    th, c = sweep(gpu_ctx, 9001, f, complex(ni ** 2, 0), complex(nt ** 2, 0))
    g0 = (nt - ni) / (nt + ni)
    assert abs(c["rTM"][0] - g0) < 1e-12 and abs(c["rTE"][0] - g0) < 1e-12
    tb = np.arctan(nt / ni)
    assert abs(th[np.argmin(np.abs(c["rTM"]))] - tb) <= (np.pi / 2) / 9000
    th, c = sweep(gpu_ctx, 9001, f, complex(nt ** 2, 0), complex(ni ** 2, 0))
    crit = np.arcsin(ni / nt)
    beyond = th > crit + 1e-6
    assert np.allclose(np.abs(c["rTE"][beyond]), 1.0, atol=1e-9) and np.allclose(np.abs(c["rTM"][beyond]), 1.0, atol=1e-9)
    assert np.all(np.abs(c["rTE"][th < crit - 1e-3]) < 1.0)
    # the reference's slab branch (inf_wall = 0) at width 0: no reflection, full transmission
    th, c = sweep(gpu_ctx, 31, f, 1 + 0j, complex(5.31, -0.2), width=0.0, inf_wall=0)
    assert np.allclose(c["rTM"], 0, atol=1e-15) and np.allclose(c["tTE"], 1, atol=1e-12)


def path_fields_given(gpu_ctx, res, mats, tri_mat, txcs, f, pt, antenna, surf_cs):
    p = res.paths()
    n = len(p["prx"])
    off = np.ascontiguousarray(p["poff"], np.int64)
    cnt = np.ascontiguousarray(p["pnseg"], np.int32)
    stype = np.ascontiguousarray(p["stype"], np.int32)
    stri = p["stri"]
    mat = np.zeros(len(stri), np.int32)
    for k in range(n):
        for s in range(1, cnt[k]):
            q = off[k] + s
            mat[q] = tri_mat[stri[q - 1]]
    params = (C.c_double * 16)()
    buf = np.zeros(16)
    buf[:9] = np.asarray(txcs, float).reshape(-1)
    buf[9], buf[10] = f, pt
    raw = bytearray(buf.tobytes())
    raw[88:92] = np.int32(1).tobytes()          # useFresnel
    raw[92:96] = np.int32(antenna).tobytes()    # antennaType
    raw[96:104] = np.float64(0.0).tobytes()     # wallWidth
    raw[104:112] = np.zeros(2, np.int32).tobytes()
    C.memmove(params, bytes(raw), 112)
    out = np.zeros((n, 3, 2))
    m = np.ascontiguousarray(mats, np.float64)
    src, tgt = np.ascontiguousarray(p["ssrc"]), np.ascontiguousarray(p["stgt"])
    surf = np.ascontiguousarray(stri, np.int32)
    cs = np.ascontiguousarray(surf_cs, np.float64)
    rc = lib().sbr_path_fields(gpu_ctx.h, C.addressof(params), m.ctypes.data, m.shape[0], n, off.ctypes.data,
                               cnt.ctypes.data, src.ctypes.data, tgt.ctypes.data, surf.ctypes.data, stype.ctypes.data,
                               mat.ctypes.data, cs.shape[0], cs.ctypes.data, out.ctypes.data)
    assert rc == 0, lib().sbr_last_error(gpu_ctx.h)
    return out[..., 0] + 1j * out[..., 1]


def test_path_fields_entry_equals_trace_fields(gpu_ctx):
    """The kept paths of a trace, handed back through sbr_path_fields with the mesh vector's frames
    (Q1 indexing), give the trace's per-path fields bit for bit."""
    say("sbr_path_fields")
    cfg = scenes.config2(tess=4)
    gpu_ctx.scene_from_soup(*cfg["soup"])
    gpu_ctx.materials(cfg["materials"])
    tr = dict(cfg["trace"], tess=4)
    res = gpu_ctx.trace(make_params(**tr), cfg["rx"])
    want = res.path_fields()
    t = gpu_ctx.scene_triangles()
    sc = O.Scene.from_soup(*cfg["soup"])
    ot = sc.triangles()
    # frames of mesh[id]: the surface at position id of the BVH-sorted vector (Q1)
    def frame(v, n):   # Triangle::findCoordSys with the library's operation order
        d = [float(v[3 + a] - v[a]) for a in range(3)]
        s = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        y = [d[0] / s, d[1] / s, d[2] / s]
        x = [y[1] * n[2] - y[2] * n[1], y[2] * n[0] - y[0] * n[2], y[0] * n[1] - y[1] * n[0]]
        return np.array(x + y + list(n))
    mesh = t["mesh_tri_id"]
    assert np.array_equal(mesh, ot["mesh_tri_id"])
    surf_cs = np.array([frame(t["verts"][mesh[i]], t["normals"][mesh[i]]) for i in range(len(mesh))])
    got = path_fields_given(gpu_ctx, res, cfg["materials"], t["material"], tr.get("txCoordSys", np.eye(3)),
                            tr["frequency"], tr.get("pt", 1.0), tr.get("antennaType", 1), surf_cs)
    assert len(want) > 100
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    assert not bad.any(), np.flatnonzero(bad.any(axis=1))[:5]
