"""Runs the C++ drop-in test program (tests/cpp/test_fastfft_api.cpp, built by
build()) on the GPU: the reference's known-answer tests written against the C++
FourierTransformer header, in FP32 and with __half I/O."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_bin", "test_fastfft_api")


@pytest.mark.parametrize("group", ["round_trip", "unit_impulse", "convolution", "scale", "half_io"])
def test_cpp_known_answers(group):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    print(f"[gpu-case] C++ drop-in test group {group}", flush=True)
    r = subprocess.run([BIN, group], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
