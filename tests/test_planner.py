"""CPU tests of the GPU-free planner and the C ABI library (no GPU calls).

Covers: the library loads and exports every symbol include/fastfft/fastfft.h
declares; the reference's size rules (ReturnPaddedMemorySize,
include/FastFFT.h:363-404) and dimension validation (ValidateDimensions,
src/fastfft/FastFFT.cu:803-930) as status codes; the stage lists per
operation; and a bounds proof: for every shape the GPU tests and the benchmark
run, every address a planned stage reads or writes lies inside the buffer the
caller allocates by the reference's sizing rules (or the plan's workspace).
"""
import ctypes
import re

import pytest

import fastfft_amd as F


def plan(fin, fout, iin, iout, origin=(0, 0, 0), rank=2, input_type="f32"):
    ft = F.FourierTransformer(input_type, "c64", rank)
    ft.SetForwardFFTPlan(*fin, *fout)
    ft.SetInverseFFTPlan(*iin, *iout)
    if any(origin):
        ft.SetOutputWindow(*origin)
    return ft


def test_library_exports_every_declared_symbol():
    L = F.lib()
    names = F.exported_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.fastfft_version()


def test_memory_sizes_and_dims():
    ft = plan((384, 384, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1))
    assert ft.ReturnInputMemorySize() == 386 * 384
    assert ft.ReturnFwdOutputMemorySize() == 2049 * 2 * 4096
    assert ft.ReturnInvOutputMemorySize() == 514 * 512
    assert ft.ReturnFwdInputDimensions() == (384, 384, 1, 193)
    assert ft.ReturnFwdOutputDimensions() == (4096, 4096, 1, 2049)
    assert ft.ReturnInvOutputDimensions() == (512, 512, 1, 257)
    ft = plan((64, 64, 1), (64, 64, 1), (64, 64, 1), (64, 64, 1))
    # src/tests/non_cuda_compilation_unit.cu:28-49: (64+2)*64 floats in and out
    assert ft.ReturnInputMemorySize() == ft.ReturnInvOutputMemorySize() == 66 * 64


@pytest.mark.parametrize("fin,fout,iin,iout,status", [
    ((64, 64, 1), (128, 128, 1), (256, 256, 1), (256, 256, 1), F.INVALID_DIMENSION),  # fwd out != inv in
    ((64, 128, 1), (128, 64, 1), (128, 64, 1), (128, 64, 1), F.INVALID_DIMENSION),  # mixed fwd change
    ((100, 100, 1), (100, 100, 1), (100, 100, 1), (100, 100, 1), F.INVALID_DIMENSION),  # non-pow2 without pad
    ((64, 64, 1), (96, 96, 1), (96, 96, 1), (96, 96, 1), F.INVALID_DIMENSION),  # non-pow2 fwd output
    ((64, 64, 1), (64, 64, 1), (64, 64, 1), (100, 100, 1), F.INVALID_DIMENSION),  # non-pow2 inv out with increase
    ((64, 64, 64), (1024, 1024, 1024), (1024, 1024, 1024), (1024, 1024, 1024), F.INVALID_DIMENSION),  # 3d > 512
    ((64, 1, 1), (64, 1, 1), (64, 1, 1), (64, 1, 1), F.UNSUPPORTED),  # 1d
    ((64, 64, 1), (16384, 16384, 1), (16384, 16384, 1), (16384, 16384, 1), F.INVALID_DIMENSION),  # too big
])
def test_validation_errors(fin, fout, iin, iout, status):
    ft = plan(fin, fout, iin, iout)
    with pytest.raises(F.FastFFTError) as e:
        ft.Describe("fwd")
    assert e.value.status == status


def test_not_planned_and_bad_ops():
    ft = F.FourierTransformer()
    ft.SetForwardFFTPlan(64, 64, 1, 64, 64, 1)
    with pytest.raises(F.FastFFTError) as e:
        ft.Describe("fwd")
    assert e.value.status == F.NOT_PLANNED
    with pytest.raises(F.FastFFTError) as e:
        ft.FwdFFT(0)
    assert e.value.status == F.INVALID_ARGUMENT  # null device pointer, caught before any GPU work
    ft.SetInverseFFTPlan(64, 64, 1, 64, 64, 1)
    with pytest.raises(F.FastFFTError) as e:
        ft.FwdImageInvFFT(16, 16, None, intra_op=F.Noop())
    assert e.value.status == F.INVALID_ARGUMENT  # FwdImageInvFFT needs an intra op
    assert F.lib().fastfft_create(None, 0, 0, 2) == F.INVALID_ARGUMENT
    with pytest.raises(F.FastFFTError) as e:
        F.FourierTransformer(rank=4)
    assert e.value.status == F.INVALID_ARGUMENT


def test_padded_flag_must_match():
    ft = F.FourierTransformer()
    ft.SetForwardFFTPlan(64, 64, 1, 64, 64, 1, True)
    with pytest.raises(F.FastFFTError):
        ft.SetInverseFFTPlan(64, 64, 1, 64, 64, 1, False)


def test_stage_lists_xcorr_4096_384_512():
    ft = plan((384, 384, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), origin=(-256, -256, 0))
    d = ft.Describe("xcorr").splitlines()[1:]
    assert [s.split()[0] for s in d] == ["fwd_real", "xcorr", "inv_c2r"]
    assert "N=4096 P=512 Q=8 L=384 M=512 o=-256" in d[1]
    assert "lines=2049" in d[1]
    assert "lines=512" in d[2] and "o=-256" in d[2]


_STAGE = re.compile(r"^(\w+) ref=(\S+) N=(\d+) P=(\d+) Q=(\d+) L=(\d+) M=(\d+) o=(-?\d+) nout=(\d+) lines=(\d+) "
                    r"(\w+)->(\w+) im=\((-?\d+),(-?\d+),(-?\d+),(\d+)\) om=\((-?\d+),(-?\d+),(-?\d+),(\d+)\) "
                    r"gm=\((-?\d+),(-?\d+),(-?\d+),(\d+)\)")


def _extent(s1, s2, es, n2, nlines, nelem):
    last = nlines - 1
    bases = [(last // n2) * s1 + (last % n2) * s2]
    if n2 > 1:
        bases.append((last // n2) * s1 + (n2 - 1) * s2)
    return max(bases) + (nelem - 1) * es + 1


def _check_bounds(ft, op, rank=2):
    a_in, fwd_out, inv_out = ft.ReturnInputMemorySize(), ft.ReturnFwdOutputMemorySize(), ft.ReturnInvOutputMemorySize()
    ws = ft.WorkspaceBytes() // 8
    cap = {  # capacity in elements of the buffer's own type
        ("fwd", "in", "r"): a_in, ("fwd", "out", "w"): fwd_out // 2,
        ("inv", "in", "r"): fwd_out // 2, ("inv", "out", "w"): inv_out,
        ("xcorr", "in", "r"): a_in, ("xcorr", "out", "w"): inv_out,
    }
    for line in ft.Describe(op).splitlines()[1:]:
        m = _STAGE.match(line)
        assert m, line
        kind = m.group(1)
        N, P, Q, L, M, o, nout, nlines = (int(m.group(i)) for i in range(3, 11))
        src, dst = m.group(11), m.group(12)
        im = [int(m.group(i)) for i in range(13, 17)]
        om = [int(m.group(i)) for i in range(17, 21)]
        assert P * Q == N and P >= max(L, M) and P <= N
        nread = {"fwd_real": L, "fwd_cpx": L, "inv_cpx": N, "inv_c2r": N // 2 + 1, "xcorr": L}[kind]
        nwrite = {"fwd_real": nout, "fwd_cpx": nout, "inv_cpx": M, "inv_c2r": M, "xcorr": M}[kind]
        rd = _extent(*im, nlines, nread)
        wr = _extent(*om, nlines, nwrite)
        if src in ("ws_a", "ws_b"):
            assert rd <= ws, line
        else:
            assert rd <= cap[(op, src, "r")], line
        if dst in ("ws_a", "ws_b"):
            assert wr <= ws, line
        else:
            assert wr <= cap[(op, dst, "w")], line
        if kind == "xcorr":
            # image_to_search C[kx][ky] (C[kz][kx][ky]) has the fwd output size; all N values of a line are read
            gm = [int(m.group(i)) for i in range(21, 25)]
            assert _extent(*gm, nlines, N) <= fwd_out // 2, line


BOUND_CASES = [
    ((64, 64, 1), (512, 512, 1), (512, 512, 1), (512, 512, 1), (0, 0, 0)),
    ((384, 384, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), (-256, -256, 0)),
    ((4096, 4096, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), (-256, -256, 0)),
    ((128, 128, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), (-256, -256, 0)),
    ((6, 5, 1), (8, 8, 1), (8, 8, 1), (4, 6, 1), (-2, -3, 0)),
    ((100, 60, 1), (256, 128, 1), (256, 128, 1), (100, 60, 1), (-50, -30, 0)),
    ((256, 256, 1), (64, 32, 1), (64, 32, 1), (64, 32, 1), (0, 0, 0)),
    ((8192, 16, 1), (8192, 16, 1), (8192, 16, 1), (8192, 16, 1), (0, 0, 0)),
    ((16, 8192, 1), (32, 8192, 1), (32, 8192, 1), (32, 300, 1), (0, 7, 0)),
    ((1000, 900, 1), (1024, 1024, 1), (1024, 1024, 1), (1000, 1000, 1), (3, -5, 0)),
    ((100, 50, 1), (2048, 512, 1), (2048, 512, 1), (300, 100, 1), (-150, -50, 0)),
    ((512, 512, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), (-256, -256, 0)),
]


@pytest.mark.parametrize("fin,fout,iin,iout,origin", BOUND_CASES)
def test_bounds_2d(fin, fout, iin, iout, origin):
    ft = plan(fin, fout, iin, iout, origin)
    for op in ("fwd", "inv", "xcorr"):
        try:
            _check_bounds(ft, op)
        except F.FastFFTError as e:
            assert e.status == F.UNSUPPORTED, (op, str(e))


BOUND_CASES_3D = [
    ((64, 64, 64), (256, 256, 256), (256, 256, 256), (256, 256, 256), (0, 0, 0)),
    ((512, 512, 512), (512, 512, 512), (512, 512, 512), (128, 128, 128), (-64, -64, -64)),
    ((32, 32, 32), (64, 64, 64), (64, 64, 64), (32, 16, 8), (-16, 0, -4)),
    ((6, 10, 12), (32, 32, 16), (32, 32, 16), (32, 32, 16), (0, 0, 0)),
]


@pytest.mark.parametrize("fin,fout,iin,iout,origin", BOUND_CASES_3D)
def test_bounds_3d(fin, fout, iin, iout, origin):
    ft = plan(fin, fout, iin, iout, origin, rank=3)
    for op in ("fwd", "inv", "xcorr"):
        try:
            _check_bounds(ft, op, rank=3)
        except F.FastFFTError as e:
            assert e.status == F.UNSUPPORTED, (op, str(e))


def test_workspace_is_reported_without_gpu():
    ft = plan((384, 384, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1))
    # correlation intermediates, row-major with an even leading dimension of 2050 complex floats:
    # ws_a: A[384 rows][2049 kx], ws_b: B[512 rows][2049 kx] (the last row ends at kx = 2048)
    assert ft.WorkspaceBytes() == 8 * ((383 * 2050 + 2049) + (511 * 2050 + 2049))
    assert ft.GetDeviceBufferPointer() == 0  # allocated at the first transform only


def test_set_precision_values():
    ft = plan((64, 64, 1), (256, 256, 1), (256, 256, 1), (64, 64, 1))
    ft.SetPrecision("split_bf16")
    ft.SetPrecision("fp32")
    with pytest.raises(F.FastFFTError) as e:
        ft.SetPrecision(7)
    assert e.value.status == F.INVALID_ARGUMENT


def test_noalloc_rejects_plans_beyond_static_scratch():
    """The allocation-free correlation checks its scratch need on the host, before any device call:
    an 8192^2 correlation with a 1024^2 template and a 2048^2 window needs about 100 MB, over the 32 MiB."""
    ft = F.FourierTransformer()
    ft.SetForwardFFTPlan(1024, 1024, 1, 8192, 8192, 1)
    ft.SetInverseFFTPlan(8192, 8192, 1, 2048, 2048, 1)
    with pytest.raises(F.FastFFTError) as e:
        ft.FwdImageInvFFTArgMaxNoAlloc(1, 2, 3, 4, 5)
    assert e.value.status == F.UNSUPPORTED and "33554432 bytes" in str(e.value)
    # a 3-D plan is refused as by the allocating entry point
    f3 = F.FourierTransformer("f32", "c64", 3)
    f3.SetForwardFFTPlan(16, 16, 16, 64, 64, 64)
    f3.SetInverseFFTPlan(64, 64, 64, 64, 64, 64)
    with pytest.raises(F.FastFFTError) as e:
        f3.FwdImageInvFFTArgMaxNoAlloc(1, 2, 3, 4, 5)
    assert e.value.status == F.UNSUPPORTED


def test_log_names_stage_lists():
    """FASTFFT_LOG=1 writes each operation's stage list to stderr when it is built (no GPU needed
    for planning); without it the library writes nothing."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import fastfft_amd as F\n"
            "ft = F.FourierTransformer()\n"
            "ft.SetForwardFFTPlan(384, 384, 1, 4096, 4096, 1)\n"
            "ft.SetInverseFFTPlan(4096, 4096, 1, 512, 512, 1)\n"
            "ft.SetOutputWindow(-256, -256)\n"
            "ft.Describe('xcorr')\n"
            "ft.Describe('xcorr')\n"
            "ft.Describe('fwd')\n")

    def run(flag):
        env = dict(os.environ, FASTFFT_LOG=flag)
        r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return [l for l in r.stderr.splitlines() if l.startswith("fastfft: ") or l.startswith("fwd") or
                l.startswith("xcorr") or l.startswith("inv")]

    lines = run("1")
    heads = [l for l in lines if l.startswith("fastfft: plan ")]
    # built once per operation: the second Describe('xcorr') reuses the list
    assert heads == ["fastfft: plan xcorr stages:", "fastfft: plan fwd stages:"]
    i = lines.index("fastfft: plan xcorr stages:")
    kinds = [l.split()[0] for l in lines[i + 1:i + 4]]
    assert kinds == ["fwd_real", "xcorr", "inv_c2r"]
    assert "N=4096 P=512 Q=8 L=384 M=512 o=-256" in lines[i + 2]
    assert run("0") == []
