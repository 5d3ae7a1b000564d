"""The CPU restatement under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/sanitize_driver.c runs every entry point of oracle/jbgrav_oracle.c on seeded
inputs and tree edge cases (one particle, a pair 1e-14 apart, coincident particles,
targets outside the root box, theta 0 and 1e30); here it is built with gcc and the
sanitizers and must finish cleanly.  Host code only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "orc_san")
    build = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-fopenmp", "-ffp-contract=off", "-o", exe,
             os.path.join(ROOT, "oracle", "sanitize_driver.c"), "-lm"]
    r = subprocess.run(build, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr, r.stderr
    assert r.stdout.strip().endswith("ok"), r.stdout
