"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code that runs without a device
(g++ only; no GPU code is built with a sanitizer):

* the FP64 oracle (oracle/sbr_oracle.cpp) driven by tests/sanitize_oracle.cpp over the reference's
  STL scenes and a synthetic box room: traces in both reception models, every physical-mode switch,
  the tabulated pattern, launch strides, tree dumps, closest hits (BVH and brute force), the
  icosahedron and the Fresnel sweep;
* the host library parts (sbrsolver_amd/csrc/host/preprocessor.cpp, ray.cpp, include/sbrsolver/BVH.h)
  driven by tests/sanitize_host.cpp: STL reading (ASCII and binary) with the Edges / Faces writers,
  the SAH build and closest hits, BVH.vtk, soups, receiver and pattern files, path trees.

Any sanitizer report ends the program with a non-zero status (-fno-sanitize-recover=all), which
fails the test with the report."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SAN = ["-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_oracle")
    subprocess.check_call(["g++"] + SAN + ["-fopenmp", os.path.join(ROOT, "oracle", "sbr_oracle.cpp"),
                                           os.path.join(ROOT, "tests", "sanitize_oracle.cpp"), "-o", exe])
    out = _run([exe, os.path.join(GOLDEN, "stl")], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    assert "box_room" in lines and lines["ico"] == ["162"]
    for name in ("ground.stl", "office.stl", "menger_sponge.stl", "box_room"):
        ntri, nnodes, work, agree = map(int, lines[name])
        assert ntri > 0 and work > 0 and agree == 200   # BVH closest hit == brute force


def test_host_library_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    host = os.path.join(ROOT, "sbrsolver_amd", "csrc", "host")
    subprocess.check_call(["g++"] + SAN + ["-I", os.path.join(ROOT, "include"), os.path.join(host, "preprocessor.cpp"),
                                           os.path.join(host, "ray.cpp"),
                                           os.path.join(ROOT, "tests", "sanitize_host.cpp"), "-o", exe])
    out_dir = tmp_path / "out"
    out_dir.mkdir()
    out = _run([exe, os.path.join(GOLDEN, "stl"), os.path.join(GOLDEN, "rx"), str(out_dir)],
               env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    assert lines["office.stl"][:2] == ["28", "9"] and lines["soup"] == ["1", "2"]
    assert lines["pattern"] == ["1", "20"] and lines["RX_Office.dat"] == ["1", "108"]
    with open(out_dir / "Faces.vtk", "rb") as a:   # written for ground.stl
        with open(os.path.join(GOLDEN, "output", "Faces.vtk"), "rb") as b:
            assert a.read() == b.read()
