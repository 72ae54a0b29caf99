"""GPU parity of the 2-D FourierTransformer path against the FP64 oracle.

Runs through the C ABI (ctypes binding) on cuda:0.  Tolerances: relative L2
error <= 2e-6 for FP32 against the FP64 oracle (the stated FP32 bar), exact
equality for the reference's closed-form known-answer tests (the reference
itself compares with == and != 0.0f, src/tests/helper_functions.cuh:91-109),
and exact argmax agreement for correlations.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_helpers as H

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 2e-6


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rand_img(x, y, seed):
    return np.random.default_rng(seed).standard_normal((y, x)).astype(np.float32)


def run_fwd(ft, host_padded, nfloats):
    buf = H.dev_zeros(nfloats)
    buf[: host_padded.size] = H.to_dev(host_padded.reshape(-1))
    ft.FwdFFT(buf)
    return buf


# ------------------------------------------------------------------ round trip 64^2 into 512^2
def test_round_trip_64_into_512():
    H.log("round trip 64x64 -> 512x512 fwd, 512 inv")
    img = rand_img(64, 64, seed=1)
    ft = H.make_plan((64, 64, 1), (512, 512, 1), (512, 512, 1), (512, 512, 1))
    n = max(ft.ReturnInputMemorySize(), ft.ReturnFwdOutputMemorySize(), ft.ReturnInvOutputMemorySize())
    buf = run_fwd(ft, H.padded_real(img, 64), n)
    spec = H.complex_from_f32(H.to_host(buf), (257, 512))
    want = O.fwd_2d(H.padded_real(img, 64), 64, 512, 512)
    assert O.rel_l2(spec, want) < TOL
    ft.InvFFT(buf)
    back = H.real_window(H.to_host(buf), 512, 512)
    want_back = O.inv_2d(H.complex_from_f32(H.complex_to_f32(spec), (257, 512)), 512, 512, 512, 512)
    assert O.rel_l2(back, want_back) < TOL
    padded = np.zeros((512, 512))
    padded[:64, :64] = img
    assert O.rel_l2(back / (512 * 512), padded) < TOL


# ------------------------------------------------------------------ golden fixtures through the GPU
def test_golden_forward(golden):
    for key in ("bf", "np"):
        in_x, in_y, n_x, n_y = (int(v) for v in golden[f"{key}_fwd_dims"])
        H.log(f"golden fwd {key} {in_x}x{in_y} -> {n_x}x{n_y}")
        ft = H.make_plan((in_x, in_y, 1), (n_x, n_y, 1), (n_x, n_y, 1), (n_x, n_y, 1))
        n = max(ft.ReturnInputMemorySize(), ft.ReturnFwdOutputMemorySize())
        buf = run_fwd(ft, golden[f"{key}_fwd_in"], n)
        got = H.complex_from_f32(H.to_host(buf), (n_x // 2 + 1, n_y))
        assert O.rel_l2(got, golden[f"{key}_fwd_out"]) < TOL


def test_golden_inverse(golden):
    for key in ("bf", "np"):
        n_x, n_y, out_x, out_y, ox, oy = (int(v) for v in golden[f"{key}_inv_dims"])
        H.log(f"golden inv {key} {n_x}x{n_y} -> {out_x}x{out_y} at ({ox},{oy})")
        ft = H.make_plan((n_x, n_y, 1), (n_x, n_y, 1), (n_x, n_y, 1), (out_x, out_y, 1), origin=(ox, oy, 0))
        spec = golden[f"{key}_inv_in"]
        inp = H.to_dev(H.complex_to_f32(spec))
        out = H.dev_poison(ft.ReturnInvOutputMemorySize())
        ft.InvFFT(inp, out)
        got = H.real_window(H.to_host(out), out_x, out_y)
        want = golden[f"{key}_inv_out"]
        # the fixture was made from the FP64 spectrum; compare with the oracle on the FP32-rounded input too
        want32 = O.inv_2d(spec.astype(np.complex64), n_x, n_y, out_x, out_y, ox, oy)
        assert O.rel_l2(got, want32) < TOL
        assert O.rel_l2(got, want) < 1e-5


def test_golden_xcorr(golden):
    tx, ty, n_x, n_y, out_x, out_y, ox, oy = (int(v) for v in golden["xc_dims"])
    H.log("golden xcorr")
    ft = H.make_plan((tx, ty, 1), (n_x, n_y, 1), (n_x, n_y, 1), (out_x, out_y, 1), origin=(ox, oy, 0))
    t = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tin = golden["xc_tmpl"].reshape(-1)
    t[: tin.size] = H.to_dev(tin)
    img = H.to_dev(H.complex_to_f32(golden["xc_image_spec"]))
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(t, img, out)
    got = H.real_window(H.to_host(out), out_x, out_y)
    want32 = O.xcorr_2d(golden["xc_tmpl"], tx, golden["xc_image_spec"].astype(np.complex64), n_x, n_y, out_x, out_y, ox, oy)
    assert O.rel_l2(got, want32) < TOL
    assert O.rel_l2(got, golden["xc_out"]) < 1e-5


# ------------------------------------------------------------------ the reference's known-answer tests
SIZES = [32, 64, 128, 256, 512, 1024, 2048, 4096]  # src/tests/tests.h:11


@pytest.mark.parametrize("inplace", [True, False])
def test_kat_unit_impulse_increase(inplace):
    # src/tests/unit_impulse_test.cu:4-218 (increase, in place and out of place)
    for i, small in enumerate(SIZES[:-1]):
        for big in SIZES[i + 1:]:
            H.log(f"unit impulse {small} -> {big} inplace={inplace}")
            ft = H.make_plan((small, small, 1), (big, big, 1), (big, big, 1), (big, big, 1))
            n = max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize(), ft.ReturnFwdOutputMemorySize())
            img = np.zeros((small, small), np.float32)
            img[0, 0] = 1.0
            a = H.dev_zeros(n)
            a[: small * H.F.padded_pitch(small)] = H.to_dev(H.padded_real(img, small).reshape(-1))
            if inplace:
                ft.FwdFFT(a)
                spec_buf = a
            else:
                b = H.dev_zeros(n)
                ft.FwdFFT(a, b)
                a.zero_()
                spec_buf = b
            spec = H.complex_from_f32(H.to_host(spec_buf), (big // 2 + 1, big))
            assert np.sum(np.abs(spec)) == (big // 2 + 1) * big
            if inplace:
                ft.InvFFT(a)
                out = a
            else:
                ft.InvFFT(spec_buf, a)
                out = a
            back = H.real_window(H.to_host(out), big, big)
            assert back.sum() == big * big
            assert back[0, 0] == big * big


def test_kat_unit_impulse_decrease():
    # unit_impulse_test.cu:29-38 with do_increase_size = false: Fourier cropping
    for i, small in enumerate(SIZES[:-1]):
        for big in SIZES[i + 1:]:
            H.log(f"unit impulse decrease {big} -> {small}")
            ft = H.make_plan((big, big, 1), (small, small, 1), (small, small, 1), (small, small, 1))
            n = max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize(), ft.ReturnFwdOutputMemorySize())
            img = np.zeros((big, big), np.float32)
            img[0, 0] = 1.0
            a = H.dev_zeros(n)
            a[:] = H.to_dev(H.padded_real(img, big).reshape(-1))
            ft.FwdFFT(a)
            spec = H.complex_from_f32(H.to_host(a), (small // 2 + 1, small))
            assert np.sum(np.abs(spec)) == (small // 2 + 1) * small
            ft.InvFFT(a)
            back = H.real_window(H.to_host(a), small, small)
            assert back.sum() == small * small


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_kat_constant_image(n):
    # src/tests/constant_image_test.cu:87-101, :186-192
    H.log(f"constant image {n}")
    ft = H.make_plan((n, n, 1), (n, n, 1), (n, n, 1), (n, n, 1))
    a = H.dev_zeros(ft.ReturnInputMemorySize())
    a[:] = H.to_dev(H.padded_real(np.ones((n, n), np.float32), n).reshape(-1))
    ft.FwdFFT(a)
    spec = H.complex_from_f32(H.to_host(a), (n // 2 + 1, n))
    assert spec[0, 0] == n * n
    spec[0, 0] = 0
    assert np.all(spec == 0)
    ft.InvFFT(a)
    back = H.real_window(H.to_host(a), n, n)
    assert back.sum() == float(n * n) * (n * n)


CONV_PAIRS = [(16, 64), (64, 64), (64, 512), (128, 1024), (256, 4096), (512, 4096), (384, 4096)]


@pytest.mark.parametrize("small,big", CONV_PAIRS)
def test_kat_impulse_convolution(small, big):
    # src/tests/padded_convolution_FastFFT_vs_cuFFT.cu:179-265: 2 (*) 3 -> N*6 at the origin, exact zeros elsewhere
    H.log(f"impulse convolution {small} -> {big}")
    ft = H.make_plan((small, small, 1), (big, big, 1), (big, big, 1), (big, big, 1))
    tgt = H.make_plan((big, big, 1), (big, big, 1), (big, big, 1), (big, big, 1))
    t = np.zeros((big, big), np.float32)
    t[0, 0] = 3.0
    tb = H.dev_zeros(tgt.ReturnFwdOutputMemorySize())
    tb[:] = H.to_dev(H.padded_real(t, big).reshape(-1))
    tgt.FwdFFT(tb)
    x = np.zeros((small, small), np.float32)
    x[0, 0] = 2.0
    n = max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize())
    a = H.dev_zeros(n)
    a[: small * H.F.padded_pitch(small)] = H.to_dev(H.padded_real(x, small).reshape(-1))
    ft.FwdImageInvFFT(a, tb, a, intra_op=H.F.ConjMul())
    out = H.real_window(H.to_host(a), big, big)
    assert out[0, 0] == big * big * 6.0
    out[0, 0] = 0
    assert np.all(out == 0.0)
    # SCALE pre-op and CONJ_MUL_THEN_SCALE give exactly 1.0 (preop_functor_scale_test.cu:229-364)
    s = 1.0 / (big * big * 6.0)
    for pre, intra in ((H.F.Scale(s), H.F.ConjMul()), (H.F.Noop(), H.F.ConjMulThenScale(s))):
        a.zero_()
        a[: small * H.F.padded_pitch(small)] = H.to_dev(H.padded_real(x, small).reshape(-1))
        ft.FwdImageInvFFT(a, tb, a, pre_op=pre, intra_op=intra)
        out = H.real_window(H.to_host(a), big, big)
        assert out[0, 0] == 1.0


# ------------------------------------------------------------------ random-data parity sweep
FWD_CASES = [((64, 64), (512, 512)), ((384, 384), (4096, 4096)), ((100, 60), (256, 128)), ((512, 512), (512, 512)),
             ((4096, 4096), (4096, 4096)), ((8192, 16), (8192, 16)), ((16, 8192), (32, 8192)), ((1000, 900), (1024, 1024)),
             ((256, 256), (64, 32)),
             # packed R2C rows: odd non-zero length on 512-point rows, a decrease from 1024 keeping 129 outputs
             ((301, 50), (512, 64)), ((1024, 64), (256, 32)),
             # full-length real rows as pairs: two pairs per workgroup (1024- and 2048-point rows), and an odd
             # pair count (one pair per workgroup)
             ((1024, 32), (1024, 32)), ((2048, 16), (2048, 16)), ((2048, 2), (2048, 2))]


@pytest.mark.parametrize("mode", ["default", "no_packed", "no_pair", "pair_group1", "pair_group2", "generic"])
@pytest.mark.parametrize("fin,fout", FWD_CASES)
def test_fwd_random(fin, fout, mode, monkeypatch):
    """Forward transforms against the FP64 oracle with the default kernels (512- and 1024-point real
    rows on the packed one-wave-per-row kernel, full-length real rows of 1024 points and up as pairs of
    rows transformed as one complex line, four pairs per workgroup for 4096-point rows and two for the
    others where the count allows), without the packed rows, with one full-length row per workgroup,
    with one and with at most two pairs per workgroup, and generic only."""
    if mode.startswith("pair_group"):
        monkeypatch.setenv("FASTFFT_PAIR_GROUP", mode[-1])
    if mode == "generic":
        monkeypatch.setenv("FASTFFT_NO_WAVE_ROWS", "1")
    if mode == "no_packed":
        monkeypatch.setenv("FASTFFT_NO_PACKED_ROWS", "1")
    if mode == "no_pair":
        monkeypatch.setenv("FASTFFT_NO_PAIR_LINES", "1")
    H.log(f"fwd random {fin} -> {fout} mode={mode}")
    img = rand_img(fin[0], fin[1], seed=fin[0] + 7 * fin[1])
    ft = H.make_plan((*fin, 1), (*fout, 1), (*fout, 1), (*fout, 1))
    n = max(ft.ReturnInputMemorySize(), ft.ReturnFwdOutputMemorySize())
    buf = run_fwd(ft, H.padded_real(img, fin[0]), n)
    got = H.complex_from_f32(H.to_host(buf), (fout[0] // 2 + 1, fout[1]))
    want = O.fwd_2d(H.padded_real(img, fin[0]), fin[0], fout[0], fout[1])
    assert O.rel_l2(got, want) < TOL


INV_CASES = [((512, 512), (512, 512), (0, 0)), ((4096, 4096), (512, 512), (-256, -256)), ((4096, 4096), (512, 512), (0, 0)),
             ((256, 128), (100, 60), (-50, -30)), ((64, 2048), (64, 300), (0, 7)), ((8192, 32), (8192, 32), (0, 0)),
             ((1024, 1024), (1000, 1000), (3, -5)),
             # windowed inverse rows: M > P = 512, origins off the 64-lane grid, non-square windows
             ((4096, 4096), (1024, 1024), (-256, -300)), ((2048, 2048), (700, 96), (13, -77)),
             ((512, 256), (300, 256), (-7, 0)),
             # 8192-point C2R rows cut to a window off the lane grid
             ((8192, 64), (3000, 64), (-1000, 0))]


@pytest.mark.parametrize("mode", ["default", "no_window", "no_packed", "no_packed_window", "generic"])
@pytest.mark.parametrize("nin,nout,origin", INV_CASES)
def test_inv_random(nin, nout, origin, mode, monkeypatch):
    """Trimmed inverses against the FP64 oracle with the default kernel choice, without the windowed
    inverse-row kernels (the packed 2048-8192-point C2R rows and the two-wave windowed rows: the generic
    line kernel takes their shapes), without any packed C2R rows, without only the packed windowed C2R
    rows, and with the generic kernel only."""
    if mode == "no_packed_window":
        monkeypatch.setenv("FASTFFT_NO_PACKED_WINDOW", "1")
    if mode == "generic":
        monkeypatch.setenv("FASTFFT_NO_WAVE_ROWS", "1")
    if mode == "no_packed":
        monkeypatch.setenv("FASTFFT_NO_PACKED_ROWS", "1")
    if mode == "no_window":
        monkeypatch.setenv("FASTFFT_NO_WINDOW_ROWS", "1")
    H.log(f"inv random {nin} -> {nout} at {origin} mode={mode}")
    rng = np.random.default_rng(nin[0] + nout[1])
    w = nin[0] // 2 + 1
    spec = (rng.standard_normal((w, nin[1])) + 1j * rng.standard_normal((w, nin[1]))).astype(np.complex64)
    ft = H.make_plan((*nin, 1), (*nin, 1), (*nin, 1), (*nout, 1), origin=(*origin, 0))
    inp = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.InvFFT(inp, out)
    got = H.real_window(H.to_host(out), nout[0], nout[1])
    want = O.inv_2d(spec, nin[0], nin[1], nout[0], nout[1], *origin)
    assert O.rel_l2(got, want) < TOL


XC_CASES = [((64, 64), (512, 512), (512, 512), (0, 0)), ((128, 128), (1024, 1024), (256, 256), (-128, -128)),
            ((96, 80), (256, 256), (64, 128), (-32, -64)), ((512, 512), (512, 512), (512, 512), (0, 0)),
            ((512, 512), (512, 512), (128, 128), (0, 0)), ((100, 50), (2048, 512), (300, 100), (-150, -50))]


@pytest.mark.parametrize("tin,n,nout,origin", XC_CASES)
def test_xcorr_random(tin, n, nout, origin):
    H.log(f"xcorr random {tin} in {n} -> {nout} at {origin}")
    rng = np.random.default_rng(tin[0] * 3 + n[1])
    t = rng.standard_normal((tin[1], tin[0])).astype(np.float32)
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((w, n[1])) + 1j * rng.standard_normal((w, n[1]))).astype(np.complex64)
    ft = H.make_plan((*tin, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0))
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = H.padded_real(t, tin[0]).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    img = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, img, out)
    got = H.real_window(H.to_host(out), nout[0], nout[1])
    want = O.xcorr_2d(H.padded_real(t, tin[0]), tin[0], spec, n[0], n[1], nout[0], nout[1], *origin)
    assert O.rel_l2(got, want) < TOL


# the fused column kernel (P in {256, 512, 1024}, origin a multiple of 64) against the oracle, with
# the DC and Nyquist columns packed into one workgroup (default) and unpacked, and the generic kernel
# forced, so every path stays pinned; the image spectra are random, not Hermitian
WAVE_CASES = [((48, 384), (64, 4096), (32, 512), (0, -256)), ((48, 384), (64, 4096), (32, 512), (0, 0)),
              ((40, 200), (64, 2048), (24, 256), (-12, -128)), ((32, 300), (32, 8192), (32, 512), (0, -256)),
              ((24, 1000), (32, 4096), (16, 1024), (0, -512)), ((16, 130), (32, 1024), (16, 256), (0, 128))]


@pytest.mark.parametrize("mode", ["default", "no_pack", "loop_form", "generic"])
@pytest.mark.parametrize("tin,n,nout,origin", WAVE_CASES)
def test_xcorr_column_kernel(tin, n, nout, origin, mode, monkeypatch):
    """The column kernel of one correlation in its single-template form (default), without the packed DC
    and Nyquist columns, in the form that loops over templates (loop_form), and the generic kernel."""
    env = {"default": {}, "no_pack": {"FASTFFT_XCORR_PACK": "0"}, "loop_form": {"FASTFFT_XCORR_ONE": "0"},
           "generic": {"FASTFFT_NO_WAVE_XCORR": "1"}}[mode]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    H.log(f"xcorr column {tin} in {n} -> {nout} at {origin} mode={mode}")
    rng = np.random.default_rng(tin[1] * 7 + n[1])
    t = rng.standard_normal((tin[1], tin[0])).astype(np.float32)
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((w, n[1])) + 1j * rng.standard_normal((w, n[1]))).astype(np.complex64)
    ft = H.make_plan((*tin, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0))
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = H.padded_real(t, tin[0]).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    img = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, img, out, intra_op=H.F.ConjMulThenScale(0.5))
    got = H.real_window(H.to_host(out), nout[0], nout[1])
    want = O.xcorr_2d(H.padded_real(t, tin[0]), tin[0], spec, n[0], n[1], nout[0], nout[1], *origin, intra_scale=0.5)
    assert O.rel_l2(got, want) < TOL


# the one-block-per-row x-pass kernels (R2C and trimmed C2R, one wave per sub-transform) against
# the oracle, and the generic line kernels forced on the same shapes
ROW_CASES = [((384, 40), (4096, 64), (512, 32), (-256, 0)), ((200, 24), (2048, 32), (256, 16), (0, 0)),
             ((1000, 16), (4096, 32), (1024, 16), (-512, 0)), ((300, 16), (8192, 32), (512, 16), (-256, -8)),
             ((130, 20), (1024, 32), (256, 32), (128, 0)), ((4096, 16), (4096, 16), (4096, 16), (0, 0))]


@pytest.mark.parametrize("mode", ["default", "generic"])
@pytest.mark.parametrize("tin,n,nout,origin", ROW_CASES)
def test_xcorr_row_kernels(tin, n, nout, origin, mode, monkeypatch):
    if mode == "generic":
        monkeypatch.setenv("FASTFFT_NO_WAVE_ROWS", "1")
    H.log(f"xcorr rows {tin} in {n} -> {nout} at {origin} mode={mode}")
    rng = np.random.default_rng(tin[0] * 5 + n[0])
    t = rng.standard_normal((tin[1], tin[0])).astype(np.float32)
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((w, n[1])) + 1j * rng.standard_normal((w, n[1]))).astype(np.complex64)
    ft = H.make_plan((*tin, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0))
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = H.padded_real(t, tin[0]).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    img = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, img, out, pre_op=H.F.Scale(0.25), post_op=H.F.Scale(2.0))
    got = H.real_window(H.to_host(out), nout[0], nout[1])
    want = O.xcorr_2d(H.padded_real(t, tin[0]), tin[0], spec, n[0], n[1], nout[0], nout[1], *origin, pre_scale=0.25,
                      post_scale=2.0)
    assert O.rel_l2(got, want) < TOL


# ------------------------------------------------------------------ the headline correlation, full size
def test_xcorr_4096_384_trim_512_centred():
    from fastfft_amd import workloads as W

    H.log("correlation 4096^2 x 384^2 -> 512^2 centred")
    mic, tmpl, lags = W.xcorr_config()
    # micrograph spectrum on the GPU (no size change), checked against the oracle
    mft = H.make_plan((4096, 4096, 1), (4096, 4096, 1), (4096, 4096, 1), (4096, 4096, 1))
    mbuf = H.dev_zeros(mft.ReturnFwdOutputMemorySize())
    mbuf[:] = H.to_dev(W.to_padded(mic).reshape(-1))
    mft.FwdFFT(mbuf)
    mspec = H.complex_from_f32(H.to_host(mbuf), (2049, 4096))
    want_spec = O.fwd_2d(W.to_padded(mic), 4096, 4096, 4096)
    assert O.rel_l2(mspec, want_spec) < TOL
    ft = H.make_plan((384, 384, 1), (4096, 4096, 1), (4096, 4096, 1), (512, 512, 1), origin=(-256, -256, 0))
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = W.to_padded(tmpl).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, mbuf, out)
    got = H.real_window(H.to_host(out), 512, 512)
    want = O.xcorr_2d(W.to_padded(tmpl), 384, mspec, 4096, 4096, 512, 512, -256, -256)
    err = O.rel_l2(got, want)
    print(f"4096/384/512 rel_l2={err:.3e} max_abs={O.max_abs(got, want):.3e} max={np.abs(want).max():.3e}")
    assert err < TOL
    gx, gy, _ = O.argmax_2d(got)
    wx, wy, _ = O.argmax_2d(want)
    assert (gx, gy) == (wx, wy)
    assert (gx - 256, gy - 256) == lags[0]
    # the library's own argmax epilogue agrees
    px, py, pv = ft.ArgMax(out, 512, 512, H.F.padded_pitch(512))
    assert (px, py) == (gx, gy)
    # the fused argmax of the trimmed C2R pass gives the same output and the same peak
    torch, d = H.torch_dev()
    peak = torch.zeros(2, dtype=torch.int32, device=d)
    val = torch.zeros(1, dtype=torch.float32, device=d)
    out2 = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFTArgMax(a, mbuf, out2, peak, val)
    assert np.array_equal(H.to_host(out2), H.to_host(out))
    assert tuple(H.to_host(peak)) == (gx, gy) and H.to_host(val)[0] == np.float32(pv)


# ------------------------------------------------------------------ fused argmax of FwdImageInvFFTArgMax
ARGMAX_CASES = [((48, 384), (64, 4096), (32, 512), (0, -256)), ((40, 200), (64, 2048), (24, 256), (-12, -128)),
                ((100, 60), (256, 128), (100, 60), (-50, -30)), ((24, 1000), (32, 4096), (16, 1024), (0, -512)),
                # 1024- and 512-point rows: the packed C2R kernel with its per-row keys, odd origins
                ((300, 40), (1024, 64), (512, 32), (-256, 0)), ((200, 30), (512, 64), (101, 20), (-37, 5)),
                # 2048- to 8192-point rows: the packed windowed C2R kernel with its per-row keys
                ((300, 24), (4096, 32), (512, 32), (-256, 0)), ((64, 30), (2048, 64), (777, 50), (13, -5)),
                ((100, 12), (8192, 16), (2000, 16), (-1000, 0))]


def _argmax_case(tin, n, nout, origin, input_type="f32", seed=0):
    torch, d = H.torch_dev()
    rng = np.random.default_rng(seed + tin[0] * 3 + n[1])
    t = rng.standard_normal((tin[1], tin[0])).astype(np.float32)
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((w, n[1])) + 1j * rng.standard_normal((w, n[1]))).astype(np.complex64)
    ft = H.make_plan((*tin, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0), input_type=input_type)
    dt = "f16" if input_type == "f16" else "f32"
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()), dt)
    tp = H.padded_real(t, tin[0]).reshape(-1)
    a[: tp.size] = H.to_dev(tp.astype(np.float16 if dt == "f16" else np.float32))
    img = H.to_dev(H.complex_to_f32(spec))
    return torch, d, ft, a, img, dt


@pytest.mark.parametrize("mode", ["fused", "separate"])
@pytest.mark.parametrize("tin,n,nout,origin", ARGMAX_CASES)
def test_fused_argmax(tin, n, nout, origin, mode, monkeypatch):
    if mode == "separate":
        monkeypatch.setenv("FASTFFT_NO_WAVE_ROWS", "1")
    H.log(f"fused argmax {tin} in {n} -> {nout} at {origin} mode={mode}")
    torch, d, ft, a, img, dt = _argmax_case(tin, n, nout, origin)
    pitch = H.F.padded_pitch(nout[0])
    ref_out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), dt)
    ft.FwdImageInvFFT(a, img, ref_out, intra_op=H.F.ConjMulThenScale(0.5))
    rx, ry, rv = ft.ArgMax(ref_out, nout[0], nout[1], pitch)
    host = H.real_window(H.to_host(ref_out), nout[0], nout[1])
    assert (rx, ry) == tuple(int(v) for v in np.unravel_index(np.argmax(host), host.shape)[::-1])
    for rep in range(3):  # the epilogue's key and counter return to zero after every call
        out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), dt)
        peak = torch.full((2,), -1, dtype=torch.int32, device=d)
        val = torch.zeros(1, dtype=torch.float32, device=d)
        ft.FwdImageInvFFTArgMax(a, img, out, peak, val, intra_op=H.F.ConjMulThenScale(0.5))
        assert np.array_equal(H.to_host(out), H.to_host(ref_out)), rep
        assert tuple(H.to_host(peak)) == (rx, ry), rep
        assert H.to_host(val)[0] == np.float32(rv), rep


def test_fused_argmax_half_output():
    tin, n, nout, origin = (48, 384), (64, 4096), (32, 512), (0, -256)
    H.log(f"fused argmax half output {tin} in {n} -> {nout}")
    torch, d, ft, a, img, dt = _argmax_case(tin, n, nout, origin, input_type="f16")
    pitch = H.F.padded_pitch(nout[0])
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), dt)
    peak = torch.zeros(2, dtype=torch.int32, device=d)
    val = torch.zeros(1, dtype=torch.float32, device=d)
    ft.FwdImageInvFFTArgMax(a, img, out, peak, val)
    rx, ry, rv = ft.ArgMax(out, nout[0], nout[1], pitch)
    assert tuple(H.to_host(peak)) == (rx, ry) and H.to_host(val)[0] == np.float32(rv)


@pytest.mark.parametrize("tin,n,nout,origin", [((300, 40), (1024, 64), (512, 32), (-256, 0)),
                                               ((201, 30), (512, 64), (101, 20), (-37, 5)),
                                               ((300, 24), (4096, 32), (512, 32), (-256, 0)),
                                               ((64, 12), (8192, 16), (2000, 16), (-1000, 0))])
def test_packed_rows_half_io(tin, n, nout, origin):
    """FP16 templates and FP16 outputs through the packed 512- and 1024-point row kernels and the packed
    windowed C2R of 4096 and 8192 points: against the FP64 oracle within FP16 storage rounding, and the
    fused peak equal to the argmax of the stored output."""
    H.log(f"packed rows, FP16 I/O {tin} in {n} -> {nout} at {origin}")
    torch, d, ft, a, img, dt = _argmax_case(tin, n, nout, origin, input_type="f16", seed=7)
    # the FP16 template as stored, and the spectrum, before the call
    tp = H.to_host(a)[: H.F.padded_pitch(tin[0]) * tin[1]].astype(np.float32).reshape(tin[1], -1)
    spec = H.complex_from_f32(H.to_host(img), (n[0] // 2 + 1, n[1]))
    pitch = H.F.padded_pitch(nout[0])
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), dt)
    peak = torch.zeros(2, dtype=torch.int32, device=d)
    val = torch.zeros(1, dtype=torch.float32, device=d)
    # a post scale keeps the correlation inside the FP16 range
    ft.FwdImageInvFFTArgMax(a, img, out, peak, val, intra_op=H.F.ConjMulThenScale(0.5), post_op=H.F.Scale(1.0 / 1024))
    got = H.real_window(H.to_host(out).astype(np.float32), nout[0], nout[1])
    want = O.xcorr_2d(tp, tin[0], spec, n[0], n[1], nout[0], nout[1], *origin, intra_scale=0.5, post_scale=1.0 / 1024)
    assert np.isfinite(got).all() and O.rel_l2(got, want) < 2e-3
    rx, ry, rv = ft.ArgMax(out, nout[0], nout[1], pitch)
    assert tuple(H.to_host(peak)) == (rx, ry) and H.to_host(val)[0] == np.float32(rv)


def test_fused_argmax_graph_replays():
    tin, n, nout, origin = (48, 384), (64, 4096), (32, 512), (0, -256)
    H.log("fused argmax in a HIP graph, replayed with eager work in between")
    torch, d, ft, a, img, dt = _argmax_case(tin, n, nout, origin)
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), dt)
    peak = torch.zeros(2, dtype=torch.int32, device=d)
    val = torch.zeros(1, dtype=torch.float32, device=d)
    side = torch.cuda.Stream()
    ft.SetStream(side.cuda_stream)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        ft.FwdImageInvFFTArgMax(a, img, out, peak, val)  # warm-up: allocations happen outside the capture
    torch.cuda.synchronize()
    want_out, want_peak, want_val = H.to_host(out).copy(), H.to_host(peak).copy(), H.to_host(val).copy()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        ft.FwdImageInvFFTArgMax(a, img, out, peak, val)
    for rep in range(4):
        out.zero_()
        peak.fill_(-1)
        val.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(H.to_host(out), want_out), rep
        assert np.array_equal(H.to_host(peak), want_peak) and np.array_equal(H.to_host(val), want_val), rep
        ft.FwdImageInvFFTArgMax(a, img, out, peak, val)  # eager call on the same plan between replays
        torch.cuda.synchronize()
        assert np.array_equal(H.to_host(peak), want_peak), rep


def test_log_names_kernel_launches():
    """FASTFFT_LOG=1: every kernel launch of a FwdImageInvFFT is named on stderr with its template
    arguments and launch shape (run in a child process, since the flag is read once per process)."""
    import os
    import re
    import subprocess
    import sys

    H.log("FASTFFT_LOG=1 kernel launch lines of a 64 -> 512 -> 128 correlation and a batch of two")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import torch\n"
            "import fastfft_amd as F\n"
            "d = torch.device('cuda:0')\n"
            "ft = F.FourierTransformer()\n"
            "ft.SetForwardFFTPlan(64, 64, 1, 512, 512, 1)\n"
            "ft.SetInverseFFTPlan(512, 512, 1, 128, 128, 1)\n"
            "ft.SetOutputWindow(-64, -64)\n"
            "a = torch.ones(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()), device=d)\n"
            "img = torch.ones(ft.ReturnFwdOutputMemorySize(), device=d)\n"
            "out = torch.zeros(ft.ReturnInvOutputMemorySize(), device=d)\n"
            "ft.FwdImageInvFFT(a, img, out)\n"
            "torch.cuda.synchronize()\n"
            "ts = [a.clone(), a.clone()]\n"
            "outs = [torch.zeros_like(out), torch.zeros_like(out)]\n"
            "ft.SetBatch(ts, [(64, 64), (64, 64)], outs)\n"
            "ft.FwdImageInvFFTBatch(img)\n"
            "torch.cuda.synchronize()\n")
    env = dict(os.environ, FASTFFT_LOG="1")
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    launches = [l for l in r.stderr.splitlines() if l.startswith("fastfft: launch ")]
    print("\n".join(launches))
    assert "fastfft: plan xcorr stages:" in r.stderr
    pat = re.compile(r"^fastfft: launch (k_\w+)(<.*>)? grid=\(\d+,\d+,\d+\) block=\d+ lds=\d+$")
    names = []
    for l in launches:
        m = pat.match(l)
        assert m, l
        names.append(m.group(1))
    # the twiddle table once, then one kernel per stage of the three-stage correlation
    assert names[0] == "k_twiddle_init" and len(names) >= 4
    # the batch form, logged when set, and its launches (those of the tuning runs and the capture)
    assert "fastfft: plan batch templates=2" in r.stderr
    assert len(names) > 4
