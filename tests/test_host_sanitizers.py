"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

The coefficient, twiddle, dealias, wall and filter tables are built on the
host (chns_tables.cpp, plain C++).  tests/host_tables_harness.cpp drives every
builder and host-only ABI entry point on exactly-sized heap buffers for all
models, sizes 16..2048 and the optional-physics switches; built here with
g++ -fsanitize=address,undefined, any out-of-bounds access or undefined
behaviour aborts it.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_tables_clean_under_asan_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "host_tables_harness"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "ternary-chns_amd", "csrc"),
           os.path.join(ROOT, "tests", "host_tables_harness.cpp"),
           os.path.join(ROOT, "ternary-chns_amd", "csrc", "chns_tables.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if b.returncode != 0 and "asan" in (b.stderr + b.stdout).lower():
        pytest.skip("sanitizer runtime not available: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert r.stdout.strip().startswith("ok 45"), r.stdout
