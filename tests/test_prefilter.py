"""Single-precision traversal decisions (sbrsolver_amd/csrc/prefilter.h, host build) against the
reference's FP64 slab and ray-triangle tests on grazing rays: a grown-box miss never hides an FP64
pass, a shrunk-box crossing never passes an FP64 miss (of the box, or of any box enclosing it: the
wide traversal's check of a winner's leaf), and neither the triangle pre-test nor the grown boxes of
the wide tree drop an FP64 hit (tests/prefilter_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_prefilter_decisions_sound(tmp_path):
    exe = str(tmp_path / "prefilter_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           os.path.join(ROOT, "tests", "prefilter_check.cpp"), "-o", exe])
    r = subprocess.run([exe, "1000000"], capture_output=True, text=True, timeout=300)
    cases, left, rpass, rmiss, tri_hits, bad, anc = map(int, r.stdout.split()[-7:])
    assert bad == 0, r.stdout
    # not vacuous: every kind of decision occurs many times
    assert cases > 500_000 and left > 100_000 and rpass > 10_000 and rmiss > 10_000 and tri_hits > 10_000, r.stdout
    assert anc == rpass, r.stdout


def test_sphere_pretest_decisions_sound(tmp_path):
    """Single-precision reception-sphere pre-test (sbrsolver_amd/csrc/capture_prefilter.h, host
    build, one-ulp reciprocals and square roots) against the FP64 test on boundary-hugging,
    end-of-segment, long and tiny cases (tests/sphere_prefilter_check.cpp): every decided case
    agrees with the FP64 result, and both decisions are frequent."""
    exe = str(tmp_path / "sphere_prefilter_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           os.path.join(ROOT, "tests", "sphere_prefilter_check.cpp"), "-o", exe])
    r = subprocess.run([exe, "4000000"], capture_output=True, text=True, timeout=300)
    cases, miss, hit, undecided, bad = map(int, r.stdout.split()[-5:])
    assert bad == 0, r.stdout + r.stderr
    assert cases == 4_000_000 and miss > 1_000_000 and hit > 500_000 and undecided < cases // 2, r.stdout


def test_correction_cone_pretest_sound(tmp_path):
    """Single-precision pre-test of the image correction's first step (sbrsolver_amd/csrc/
    correction_prefilter.h, host build) against the FP64 Moller-Trumbore test it stands in front of,
    on receivers hugging the cone through the triangle's edges and vertices, slivers, image sources in
    or near the triangle's plane (tests/cone_prefilter_check.cpp): no receiver that the FP64 test
    passes is rejected, every accepted receiver lies inside the row / column intervals derived from
    the same planes, and the pre-test rejects most receivers that fail."""
    exe = str(tmp_path / "cone_prefilter_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           os.path.join(ROOT, "tests", "cone_prefilter_check.cpp"), "-o", exe])
    r = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=300)
    cases, receivers, rejected, passes, bad_reject, bad_interval = map(int, r.stdout.split()[-6:])
    assert bad_reject == 0 and bad_interval == 0, r.stdout
    assert receivers == 64 * cases and passes > 100_000 and rejected > 0.8 * (receivers - passes), r.stdout
