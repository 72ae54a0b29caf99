"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The planner (fastfft_amd/csrc/planner.cpp) and the FP64 oracle (oracle/fastfft_oracle.c) are built with
-fsanitize=address,undefined by g++ / gcc and driven over every shape class the tests use (tests/asan/):
any out-of-bounds access, leak or undefined operation fails the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", "-g", "-O1"]


def _build_and_run(tmp_path, compiler, sources, extra, name):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not found")
    exe = str(tmp_path / name)
    cmd = [compiler, *SAN, *extra, *sources, "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stderr}"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return out


def test_planner_under_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "g++", [os.path.join(ROOT, "tests", "asan", "planner_driver.cpp"),
                                           os.path.join(ROOT, "fastfft_amd", "csrc", "planner.cpp")],
                         ["-std=c++17"], "planner_driver")
    assert "0 problems" in out, out


def test_oracle_under_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "gcc", [os.path.join(ROOT, "tests", "asan", "oracle_driver.c"),
                                           os.path.join(ROOT, "oracle", "fastfft_oracle.c")],
                         ["-std=gnu11", "-fopenmp"], "oracle_driver")
    assert "oracle driver: ok" in out, out
