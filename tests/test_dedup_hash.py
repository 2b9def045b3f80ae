"""Duplicate-removal group hashes (sbrsolver_amd/csrc/dedup_hash.h, host build): no two distinct
(transmitter, depth, coplanar sequence, receiver) groups of a structured key set share a slot hash
or a check word.  Regression test for a seed that combined the transmitter index with a plain xor,
which a batch of transmitters turned into slot-hash collisions."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_group_hashes_unique(tmp_path):
    exe = str(tmp_path / "dedup_hash_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           os.path.join(ROOT, "tests", "dedup_hash_check.cpp"), "-o", exe])
    r = subprocess.run([exe, "48", "40", "12"], capture_output=True, text=True, timeout=120)
    n, da, db = map(int, r.stdout.split())
    assert n == 48 * (1 + 40 + 1600) * 12
    assert (da, db) == (0, 0), r.stdout
