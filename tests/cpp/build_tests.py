"""Builds the C++ drop-in test program against include/fastfft/FastFFT.h.

Plain host C++ (g++), linked to the in-tree libfastfft_amd.so and the HIP
runtime; the binary lands in tests/cpp/_bin/ (not committed) and finds the
library through its rpath, so it runs from any working directory.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "_bin", "test_fastfft_api")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build() -> str:
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
           os.path.join(HERE, "test_fastfft_api.cpp"), "-o", BIN,
           f"-L{os.path.join(ROOT, 'fastfft_amd')}", "-lfastfft_amd", f"-L{ROCM}/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../../fastfft_amd", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{' '.join(cmd)}\n{r.stderr}")
    return BIN


if __name__ == "__main__":
    print(build())
    sys.exit(0)
