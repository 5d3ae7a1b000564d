"""The octree fill's node-end search on the CPU, under AddressSanitizer and
UndefinedBehaviorSanitizer.

gravhopper_amd/csrc/gh_search.h is the code the device fill kernel runs (host and device
compile the same functions).  tests/host/end_search_check.cpp builds its search levels
from their definition in a buffer of exactly the size the library allocates, with
noise in the unused bytes, and compares each node end with the one found by comparing
octant digits; it covers 1..20 particles and nodes of more than 4096 particles (the
levels searched one at a time).  Host code only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_node_end_search_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "end_search_check")
    build = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-o", exe,
             os.path.join(ROOT, "tests", "host", "end_search_check.cpp")]
    r = subprocess.run(build, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr, r.stderr
    assert " 0 wrong" in r.stdout, r.stdout
