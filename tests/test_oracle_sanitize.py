"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host only, no GPU).

``oracle/sanitize_check.cpp`` calls every oracle function (cost, SGM and its adjoint, filter,
median, both fusion rule sets with crowded cells) on tiny and ragged shapes in FP64 and FP32;
built with ``-fsanitize=address,undefined -fno-sanitize-recover=all``, any out-of-bounds
access, leak or undefined arithmetic ends the run with a nonzero status.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "sanitize_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_check")
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-static-libasan",
         "-I", os.path.join(ROOT, "oracle"), SRC, "-o", exe],
        capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.count(": f64 ok, f32 ok") == 5
