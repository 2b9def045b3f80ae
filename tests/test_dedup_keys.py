"""Duplicate-removal keys (sbrsolver_amd/csrc/dedup_keys.h, host build): sequences interned into a
trie the way the device does, paired with receivers, give unique trie keys and unique group keys
that never equal the empty-slot markers, over a structured key set (every sequence of length <= 3
over 24 coplanar ids and both interaction types, for 16 transmitters, with 8 receivers)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_group_keys_unique(tmp_path):
    exe = str(tmp_path / "dedup_keys_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "sbrsolver_amd", "csrc"),
                           os.path.join(ROOT, "tests", "dedup_keys_check.cpp"), "-o", exe])
    r = subprocess.run([exe, "16", "24", "8"], capture_output=True, text=True, timeout=120)
    n, dt, dg, marker = map(int, r.stdout.split())
    seqs = 16 * (1 + 48 + 48 ** 2 + 48 ** 3)
    assert n == seqs * 8
    assert (dt, dg, marker) == (0, 0, 0), r.stdout
