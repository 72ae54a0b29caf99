"""Pins the FP64 CPU oracle (oracle/) before anything is compared against it.

Each oracle function is checked against something it was not derived from:
explicit DFT matrices (brute force), numpy.fft (pocketfft FP64), the committed
golden fixtures (tests/golden/make_golden.py, numpy only), a direct real-space
correlation sum, and the reference's own closed-form known-answer tests:

  constant image -> DC = N_total, zero elsewhere  (src/tests/constant_image_test.cu:87-101)
  unit impulse   -> sum |X| = #complex outputs     (src/tests/unit_impulse_test.cu:106-155)
  round trip     -> Inv(Fwd(x)) = N_total * x      (src/tests/non_cuda_compilation_unit.cu:71-83)
  impulse 2 (*) impulse 3 -> N*6 at origin, 0 else (src/tests/padded_convolution_FastFFT_vs_cuFFT.cu:179-212)
  SCALE(1/(N*6)) pre-op / CONJ_MUL_THEN_SCALE -> exactly 1.0 at origin
                                                   (src/tests/preop_functor_scale_test.cu:229-364)

The last group plants plausible mistakes (conjugation order, untransposed
layout, dropped Hermitian term, wrong window sign) in copies of oracle output
and shows that the checks above reject each of them.
"""
import numpy as np
import pytest

from oracle import oracle as O


def dftm(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64])
def test_fft1d_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.fft1d(x, -1), dftm(n, -1) @ x, atol=1e-12)
    assert np.allclose(O.fft1d(x, +1), dftm(n, +1) @ x, atol=1e-12)


def test_fft1d_numpy_large():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(4096) + 1j * rng.standard_normal(4096)
    assert O.rel_l2(O.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert O.rel_l2(O.fft1d(x, +1), np.fft.ifft(x) * 4096) < 1e-14


def test_fwd_2d_golden(golden):
    in_x, in_y, n_x, n_y = golden["bf_fwd_dims"]
    got = O.fwd_2d(golden["bf_fwd_in"], in_x, n_x, n_y)
    assert O.max_abs(got, golden["bf_fwd_out"]) < 1e-12
    in_x, in_y, n_x, n_y = golden["np_fwd_dims"]
    got = O.fwd_2d(golden["np_fwd_in"], in_x, n_x, n_y)
    assert O.rel_l2(got, golden["np_fwd_out"]) < 1e-14


def test_inv_2d_golden(golden):
    n_x, n_y, out_x, out_y, ox, oy = golden["bf_inv_dims"]
    got = O.inv_2d(golden["bf_inv_in"], n_x, n_y, out_x, out_y, ox, oy)
    assert O.max_abs(got, golden["bf_inv_out"]) < 1e-12
    n_x, n_y, out_x, out_y, ox, oy = golden["np_inv_dims"]
    got = O.inv_2d(golden["np_inv_in"], n_x, n_y, out_x, out_y, ox, oy)
    assert O.rel_l2(got, golden["np_inv_out"]) < 1e-13


def test_xcorr_2d_golden_and_direct_sum(golden):
    tx, ty, n_x, n_y, out_x, out_y, ox, oy = golden["xc_dims"]
    got = O.xcorr_2d(golden["xc_tmpl"], tx, golden["xc_image_spec"], n_x, n_y, out_x, out_y, ox, oy)
    assert O.rel_l2(got, golden["xc_out"]) < 1e-13
    # independent of any FFT: the real-space correlation sum, windowed; unnormalised
    # transforms scale it by N_total (non_cuda_compilation_unit.cu:71-83)
    direct = n_x * n_y * np.roll(golden["xc_direct_full"], (-oy, -ox), axis=(0, 1))[:out_y, :out_x]
    assert O.rel_l2(got, direct) < 1e-12


def test_3d_golden(golden):
    in_x, in_y, in_z, n_x, n_y, n_z = golden["f3_dims"]
    got = O.fwd_3d(golden["f3_in"], in_x, n_x, n_y, n_z)
    assert O.rel_l2(got, golden["f3_out"]) < 1e-14
    n_x, n_y, n_z, out_x, out_y, out_z, ox, oy, oz = golden["i3_dims"]
    got = O.inv_3d(golden["i3_in"], n_x, n_y, n_z, out_x, out_y, out_z, ox, oy, oz)
    assert O.rel_l2(got, golden["i3_out"]) < 1e-13


@pytest.mark.parametrize("in_xy,n_xy", [((64, 64), (512, 512)), ((24, 40), (64, 128)), ((30, 17), (32, 32)),
                                        ((96, 96), (128, 128)), ((48, 48), (64, 64))])
def test_fwd_2d_numpy(in_xy, n_xy):
    (in_x, in_y), (n_x, n_y) = in_xy, n_xy
    rng = np.random.default_rng(in_x * 1000 + in_y)
    img = np.zeros((in_y, O.padded_pitch(in_x)), np.float32)
    img[:, :in_x] = rng.standard_normal((in_y, in_x))
    a = np.zeros((n_y, n_x))
    a[:in_y, :in_x] = img[:, :in_x]
    assert O.rel_l2(O.fwd_2d(img, in_x, n_x, n_y), np.fft.rfft2(a).T) < 1e-14


def test_fwd_2d_decrease_keeps_first_frequencies():
    # Fourier cropping keeps k < P of the full transform (output-pruned DIT of FastFFT.cu:1232-1261)
    rng = np.random.default_rng(5)
    img = np.zeros((64, 66), np.float32)
    img[:, :64] = rng.standard_normal((64, 64))
    full = np.fft.rfft2(img[:, :64].astype(np.float64)).T  # [kx][ky], 33 x 64
    got = O.fwd_2d(img, 64, 16, 16)
    assert O.rel_l2(got, full[:9, :16]) < 1e-14


@pytest.mark.parametrize("n,out,o", [((64, 64), (64, 64), (0, 0)), ((128, 64), (32, 16), (0, 0)),
                                     ((64, 128), (16, 32), (-8, -16)), ((32, 32), (20, 12), (5, -3))])
def test_inv_2d_numpy(n, out, o):
    (n_x, n_y), (out_x, out_y), (ox, oy) = n, out, o
    rng = np.random.default_rng(n_x + 3 * n_y)
    spec = rng.standard_normal((n_x // 2 + 1, n_y)) + 1j * rng.standard_normal((n_x // 2 + 1, n_y))
    cols = np.fft.ifft(spec, axis=1) * n_y
    full = np.fft.irfft(cols.T, n=n_x, axis=1) * n_x
    want = np.roll(full, (-oy, -ox), axis=(0, 1))[:out_y, :out_x]
    assert O.rel_l2(O.inv_2d(spec, n_x, n_y, out_x, out_y, ox, oy), want) < 1e-13


def test_3d_numpy_padded_and_trimmed():
    rng = np.random.default_rng(11)
    vol = np.zeros((6, 10, O.padded_pitch(12)), np.float32)
    vol[..., :12] = rng.standard_normal((6, 10, 12))
    v = np.zeros((16, 32, 32))
    v[:6, :10, :12] = vol[..., :12]
    spec = O.fwd_3d(vol, 12, 32, 32, 16)
    assert O.rel_l2(spec, np.transpose(np.fft.rfftn(v), (0, 2, 1))) < 1e-14
    full = np.fft.irfftn(np.transpose(spec, (0, 2, 1)), s=(16, 32, 32), axes=(0, 1, 2)) * (16 * 32 * 32)
    want = np.roll(full, (4, 8, 8), axis=(0, 1, 2))[:8, :16, :16]
    got = O.inv_3d(spec, 32, 32, 16, 16, 16, 8, -8, -8, -4)
    assert O.rel_l2(got, want) < 1e-13


# ---- the reference's closed-form known-answer tests -------------------------------------------

@pytest.mark.parametrize("n", [16, 64, 256])
def test_kat_constant_image(n):
    img = np.zeros((n, O.padded_pitch(n)), np.float32)
    img[:, :n] = 1.0
    spec = O.fwd_2d(img, n, n, n)
    assert spec[0, 0] == pytest.approx(n * n)
    rest = spec.copy()
    rest[0, 0] = 0
    assert np.abs(rest).max() < 1e-9 * n * n
    back = O.inv_2d(spec, n, n, n, n)
    assert np.allclose(back, n * n, rtol=1e-12)  # round trip sum = N_total^2


@pytest.mark.parametrize("small,big", [(16, 64), (64, 512), (32, 32)])
def test_kat_unit_impulse(small, big):
    img = np.zeros((small, O.padded_pitch(small)), np.float32)
    img[0, 0] = 1.0
    spec = O.fwd_2d(img, small, big, big)
    assert np.sum(np.abs(spec)) == pytest.approx((big // 2 + 1) * big, rel=1e-12)
    back = O.inv_2d(spec, big, big, big, big)
    assert np.sum(back) == pytest.approx(big * big, rel=1e-12)
    assert back[0, 0] == pytest.approx(big * big, rel=1e-12)


@pytest.mark.parametrize("small,big", [(64, 64), (16, 128)])
def test_kat_impulse_convolution_and_scale(small, big):
    tmpl = np.zeros((small, O.padded_pitch(small)), np.float32)
    tmpl[0, 0] = 2.0
    target = np.zeros((big, O.padded_pitch(big)), np.float32)
    target[0, 0] = 3.0
    ispec = O.fwd_2d(target, big, big, big)
    out = O.xcorr_2d(tmpl, small, ispec, big, big, big, big)
    assert out[0, 0] == big * big * 6.0
    rest = out.copy()
    rest[0, 0] = 0
    assert np.abs(rest).max() < 1e-9
    s = 1.0 / (big * big * 6.0)
    assert O.xcorr_2d(tmpl, small, ispec, big, big, big, big, pre_scale=s)[0, 0] == pytest.approx(1.0, abs=1e-15)
    assert O.xcorr_2d(tmpl, small, ispec, big, big, big, big, intra_scale=s)[0, 0] == pytest.approx(1.0, abs=1e-15)


def test_correlation_peak_at_plus_shift():
    # CONJ_MUL = image * conj(template): a copy of the template at +(dx, dy) peaks at lag +(dx, dy)
    rng = np.random.default_rng(21)
    n, t = 64, 12
    tm = rng.standard_normal((t, t)).astype(np.float32)
    img = np.zeros((n, O.padded_pitch(n)), np.float32)
    dx, dy = 9, -5
    for y in range(t):
        for x in range(t):
            img[(y + dy) % n, (x + dx) % n] += tm[y, x]
    ispec = O.fwd_2d(img, n, n, n)
    tp = np.zeros((t, O.padded_pitch(t)), np.float32)
    tp[:, :t] = tm
    out = O.xcorr_2d(tp, t, ispec, n, n, 32, 32, -16, -16)
    ix, iy, _ = O.argmax_2d(out)
    assert (ix - 16, iy - 16) == (dx, dy)


@pytest.mark.parametrize("x,y,z", [(64, 64, 1), (384, 384, 1), (512, 512, 1), (4096, 4096, 1), (63, 10, 1),
                                   (64, 64, 64)])
def test_padded_memory_size_rule(x, y, z):
    # FastFFT.h:363-404: 2 * (x/2+1) * y * z floats for even x, 2 * ((x-1)/2+1) * y * z for odd x
    w = x // 2 + 1 if x % 2 == 0 else (x - 1) // 2 + 1
    assert O.padded_memory_size(x, y, z) == 2 * w * y * z


# ---- planted mistakes: the checks above must reject each of these ------------------------------

def _planted_cases(golden):
    tx, ty, n_x, n_y, out_x, out_y, ox, oy = golden["xc_dims"]
    good = O.xcorr_2d(golden["xc_tmpl"], tx, golden["xc_image_spec"], n_x, n_y, out_x, out_y, ox, oy)
    want = golden["xc_out"]
    # (a) conj on the wrong operand: template * conj(image) mirrors the correlation
    tmpl_spec = O.fwd_2d(golden["xc_tmpl"], tx, n_x, n_y)
    wrong_conj = O.inv_2d(np.conj(golden["xc_image_spec"]) * tmpl_spec, n_x, n_y, out_x, out_y, ox, oy)
    # (b) window origin sign flipped
    wrong_origin = O.xcorr_2d(golden["xc_tmpl"], tx, golden["xc_image_spec"], n_x, n_y, out_x, out_y, -ox + 1, -oy + 1)
    # (c) spectrum stored untransposed (C[ky][kx] read as C[kx][ky]) on a non-square case
    in_x, in_y, fn_x, fn_y = golden["np_fwd_dims"]
    spec = O.fwd_2d(golden["np_fwd_in"], in_x, fn_x, fn_y)
    # (d) a dropped Hermitian factor 2 in the C2R
    n_x2, n_y2, ox2, oy2, ax, ay = golden["np_inv_dims"]
    inv_good = O.inv_2d(golden["np_inv_in"], n_x2, n_y2, ox2, oy2, ax, ay)
    return good, want, wrong_conj, wrong_origin, spec, golden["np_fwd_out"], inv_good, golden["np_inv_out"]


def test_planted_mistakes_are_caught(golden):
    good, want, wrong_conj, wrong_origin, spec, spec_want, inv_good, inv_want = _planted_cases(golden)
    tol = 1e-6
    assert O.rel_l2(good, want) < tol
    assert O.rel_l2(wrong_conj, want) > 0.1
    assert O.rel_l2(wrong_origin, want) > 0.1
    # untransposed layout cannot even match shape-wise on this non-square case
    assert spec.shape == spec_want.shape and spec.T.shape != spec_want.shape
    # a half-weighted Hermitian sum (a dropped factor 2) is rejected
    assert O.rel_l2(inv_good, inv_want) < tol
    assert O.rel_l2(0.5 * inv_good, inv_want) > 0.1
    # a sign error in the forward exponent (conjugated spectrum) is rejected
    assert O.rel_l2(np.conj(spec), spec_want) > 0.1
