"""The C++ call surface of the tabulated pattern (host only, no GPU): Preprocessor::ReadPatternFile
reads a file in TXPattern.dat's layout (a 4-token header, then theta, phi, theta gain, theta phase,
phi gain, phi phase per line, theta fastest) into a GainMap, skipping lines that are not six
numbers, and RTSolver::SetTxPattern accepts a full evenly spaced grid (switching antennaType to 3)
and rejects a grid with a hole or an empty map; RTSolverOptions::SetPhysical() switches the physical
mode on."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sbrsolver_amd")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libsbr_amd.so")):
        pytest.skip("library not built")
    exe = str(tmp_path_factory.mktemp("pat") / "pattern_surface_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "pattern_surface_check.cpp"), "-o", exe, "-L", LIBDIR,
                           "-lsbr_amd", "-Wl,-rpath," + LIBDIR])
    return exe


@pytest.mark.parametrize("step", [1, 10])
def test_pattern_file_to_solver(checker, tmp_path, step):
    nt, np_ = 180 // step + 1, 360 // step
    lines = [f"{nt} {np_} {step} {step}", "# a comment line", "1 2 3 4 x 6"]
    for j in range(np_):
        for i in range(nt):
            lines.append(f"{i * step} {j * step} {-3.5 + 0.01 * i} {j * 0.5} {-40 + 0.02 * j} 0")
    path = tmp_path / "TXPattern.dat"
    path.write_text("\n".join(lines) + "\n")
    lines = subprocess.check_output([checker, str(path)], text=True).splitlines()
    ok, n, rc, typ, rc_holed, rc_empty, ok_missing = (int(x) for x in lines[0].split())
    # RTSolverOptions: reference defaults, and SetPhysical() switches every correction on (not the
    # wall thickness)
    assert lines[1].split() == ["1", "0", "0", "0", "0", "0", "0", "1", "1", "1", "1", "1", "0"]
    assert ok == 1 and n == nt * np_
    assert rc == 0 and typ == 3
    assert rc_holed < 0 and rc_empty < 0
    assert ok_missing == 0
