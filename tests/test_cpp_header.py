"""The C++ drop-in header compiles as plain host C++ (the reference's
src/tests/non_cuda_compilation_unit.cu check) and the C++ test program links
against the in-tree library.  CPU only: nothing here runs on a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def test_header_compiles_as_host_cpp(tmp_path):
    src = tmp_path / "use_header.cpp"
    src.write_text(
        '#include "fastfft/FastFFT.h"\n'
        "int main() {\n"
        "  FastFFT::KernelFunction::my_functor<float, 4, FastFFT::KernelFunction::CONJ_MUL> cm;\n"
        "  static_assert(FastFFT::HasIntraOpFunctor<decltype(cm)>);\n"
        "  FastFFT::KernelFunction::my_functor<float, 0, FastFFT::KernelFunction::NOOP> noop;\n"
        "  static_assert(!FastFFT::HasIntraOpFunctor<decltype(noop)>);\n"
        "  using FT2 = FastFFT::FourierTransformer<float, float, float2, 2>;\n"
        "  using FT3 = FastFFT::FourierTransformer<float, __half, __half2, 3>;\n"
        "  return sizeof(FT2) + sizeof(FT3) > 0 ? 0 : 1;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include",
                        f"-I{ROCM}/include", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_test_program_links():
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "cpp"))
    import build_tests

    path = build_tests.build()
    assert os.access(path, os.X_OK)
