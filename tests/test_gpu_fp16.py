"""FP16 I/O parity: FourierTransformer<float, __half, __half2, 2> through the C ABI.

Templates and micrographs are stored as FP16, the micrograph spectrum as __half2 (produced by the
library's own FwdFFT of the FP16 micrograph), the correlation output as FP16; all arithmetic is FP32.
The transforms are unnormalised, so FP16 storage needs the SCALE functors a caller would use: the
micrograph FwdFFT is pre-scaled by 1/n (spectrum values of order 1) and the correlation post-scaled by
1/(n t^2) (peak of order 0.3), both applied identically in the oracle.
The reference for each case is the FP64 oracle on the same FP16-rounded inputs, with the spectrum the
oracle computes from the FP16 micrograph, so the measured error contains the FP16 storage of the
spectrum and of the output.  Peaks must be identical to the oracle's and to the planted lag.
"""
import numpy as np
import pytest

from oracle import oracle as O

from . import gpu_helpers as H

pytestmark = pytest.mark.gpu

# FP16 storage of the spectrum and of the output: relative rounding 2^-11 per value
TOL_F16 = 2e-3


def _case(n, t, m, radius, n_templates, seed):
    from fastfft_amd import workloads as W

    torch, d = H.torch_dev()
    tmpls = [W.masked_template(t, radius, seed=seed + i).astype(np.float16) for i in range(n_templates)]
    lags = W.copy_lags(n_templates, n, m, t, seed=seed + 50)
    rng = np.random.default_rng(seed + 60)
    mic = rng.standard_normal((n, n)).astype(np.float32)
    for i, tp in enumerate(tmpls):  # template i planted inside the window at lags[i]
        lx, ly = (int(rng.integers(-m // 2 + 8, m // 2 - 8)), int(rng.integers(-m // 2 + 8, m // 2 - 8))) if i else lags[0]
        lags[i] = (lx, ly)
        ys, xs = (np.arange(t) + ly) % n, (np.arange(t) + lx) % n
        mic[np.ix_(ys, xs)] += 0.5 * tp.astype(np.float32)
    mic16 = mic.astype(np.float16)
    # micrograph spectrum on the GPU as __half2
    mft = H.make_plan((n, n, 1), (n, n, 1), (n, n, 1), (n, n, 1), input_type="f16", image_type="c32h")
    mbuf = H.to_dev(W.to_padded(mic16, dtype=np.float16).reshape(-1))
    mft.FwdFFT(mbuf, pre_op=H.F.Scale(1.0 / n))
    return torch, d, tmpls, lags, mic16, mbuf


def _spec_f64(mic16, n):
    from fastfft_amd import workloads as W

    return O.fwd_2d(W.to_padded(mic16.astype(np.float32)), n, n, n, pre_scale=1.0 / n)


@pytest.mark.parametrize("n,t,m", [(1024, 128, 256), (4096, 512, 512)])
def test_fp16_correlation_vs_oracle(n, t, m):
    from fastfft_amd import workloads as W

    H.log(f"fp16 correlation {n}^2 x {t}^2 -> {m}^2")
    torch, d, tmpls, lags, mic16, mbuf = _case(n, t, m, 0.43 * t, 1, seed=n + t)
    ft = H.make_plan((t, t, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0), input_type="f16",
                     image_type="c32h")
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()), "f16")
    tp = W.to_padded(tmpls[0], dtype=np.float16).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize(), "f16")
    peak = torch.zeros(2, dtype=torch.int32, device=d)
    val = torch.zeros(1, dtype=torch.float32, device=d)
    post = 1.0 / (n * t * t)
    ft.FwdImageInvFFTArgMax(a, mbuf, out, peak, val, post_op=H.F.Scale(post))
    got = H.real_window(H.to_host(out).astype(np.float32), m, m)
    want = O.xcorr_2d(W.to_padded(tmpls[0].astype(np.float32)), t, _spec_f64(mic16, n), n, n, m, m, -m // 2, -m // 2,
                      post_scale=post)
    err = O.rel_l2(got, want)
    print(f"fp16 {n}/{t}/{m}: rel_l2={err:.3e}")
    assert err < TOL_F16
    wx, wy, _ = O.argmax_2d(want)
    assert tuple(H.to_host(peak)) == (wx, wy)
    assert (wx - m // 2, wy - m // 2) == tuple(lags[0])


def test_fp16_batch_4096_512_templates():
    """Full size: 4096^2 FP16 micrograph, 512^2 FP16 templates (radius 220), 512^2 window; a batch
    of 4 templates, each planted once inside the window; peaks equal the planted lags."""
    from fastfft_amd import workloads as W

    n, t, m = 4096, 512, 512
    H.log("fp16 batch 4096^2 x 512^2 -> 512^2")
    torch, d, tmpls, lags, mic16, mbuf = _case(n, t, m, 220.0, 4, seed=5)
    ft = H.make_plan((t, t, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0), input_type="f16",
                     image_type="c32h")
    tb = [H.to_dev(W.to_padded(tp, dtype=np.float16).reshape(-1)) for tp in tmpls]
    outs = [H.dev_zeros(ft.ReturnInvOutputMemorySize(), "f16") for _ in tmpls]
    peaks = torch.zeros(2 * len(tmpls), dtype=torch.int32, device=d)
    vals = torch.zeros(len(tmpls), dtype=torch.float32, device=d)
    ft.SetBatch(tb, [(t, t)] * len(tmpls), outs, peaks, vals)
    post = 1.0 / (n * t * t)
    ft.FwdImageInvFFTBatch(mbuf, post_op=H.F.Scale(post))
    pk = H.to_host(peaks).reshape(-1, 2)
    for i in range(len(tmpls)):
        assert (pk[i][0] - m // 2, pk[i][1] - m // 2) == tuple(lags[i]), i
    # template 0 against the oracle
    got = H.real_window(H.to_host(outs[0]).astype(np.float32), m, m)
    want = O.xcorr_2d(W.to_padded(tmpls[0].astype(np.float32)), t, _spec_f64(mic16, n), n, n, m, m, -m // 2, -m // 2,
                      post_scale=post)
    assert O.rel_l2(got, want) < TOL_F16


def _half2_spectrum_host(buf, w, ny):
    """The __half2 spectrum C[kx][ky] of a device FP16 buffer as complex128."""
    a = H.to_host(buf)[: 2 * w * ny].astype(np.float64)
    return (a[0::2] + 1j * a[1::2]).reshape(w, ny)


@pytest.mark.parametrize("n,t,m,count", [(1024, 128, 256, 1), (4096, 384, 512, 3)])
def test_fp32_templates_against_half2_spectrum(n, t, m, count):
    """FourierTransformer<float, float, __half2, 2>: FP32 templates and outputs against a micrograph spectrum
    stored as __half2 (the cisTEM practice of keeping the image spectrum in half precision).  The oracle
    correlates the same FP32 templates with the same __half2-rounded spectrum, so the FP32 tolerance
    applies; single correlations and a batch (on the reduced frame at 4096^2) both run."""
    from fastfft_amd import workloads as W

    H.log(f"<float, float, __half2, 2> correlation {n}^2 x {t}^2 -> {m}^2, {count} templates")
    torch, d, tmpls16, lags, mic16, mbuf = _case(n, t, m, 0.43 * t, count, seed=3 * n + t)
    spec = _half2_spectrum_host(mbuf, n // 2 + 1, n)
    # the planted templates were FP16-rounded; the FP32 templates here are those same values
    tmpls = [tp.astype(np.float32) for tp in tmpls16]
    ft = H.make_plan((t, t, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0), input_type="f32",
                     image_type="c32h")
    post = 1.0 / (t * t)
    outs = []
    for tp in tmpls:
        a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
        tpp = W.to_padded(tp).reshape(-1)
        a[: tpp.size] = H.to_dev(tpp)
        out = H.dev_poison(ft.ReturnInvOutputMemorySize())
        peak = torch.zeros(2, dtype=torch.int32, device=d)
        val = torch.zeros(1, dtype=torch.float32, device=d)
        ft.FwdImageInvFFTArgMax(a, mbuf, out, peak, val, post_op=H.F.Scale(post))
        outs.append((out, peak))
    bft = H.make_plan((t, t, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0), input_type="f32",
                      image_type="c32h")
    tb = [H.to_dev(W.to_padded(tp).reshape(-1)) for tp in tmpls]
    bouts = [H.dev_poison(bft.ReturnInvOutputMemorySize()) for _ in tmpls]
    bpk = torch.full((2 * count,), -1, dtype=torch.int32, device=d)
    bval = torch.zeros(count, dtype=torch.float32, device=d)
    bft.SetBatch(tb, [(t, t)] * count, bouts, bpk, bval)
    bft.FwdImageInvFFTBatch(mbuf, post_op=H.F.Scale(post))
    bpk = H.to_host(bpk).reshape(-1, 2)
    for i, tp in enumerate(tmpls):
        want = O.xcorr_2d(W.to_padded(tp), t, spec, n, n, m, m, -m // 2, -m // 2, post_scale=post)
        wx, wy, _ = O.argmax_2d(want)
        got = H.real_window(H.to_host(outs[i][0]), m, m)
        bgot = H.real_window(H.to_host(bouts[i]), m, m)
        print(f"template {i}: single rel_l2={O.rel_l2(got, want):.3e} batch rel_l2={O.rel_l2(bgot, want):.3e}")
        assert O.rel_l2(got, want) < 2e-6, i
        assert O.rel_l2(bgot, want) < 2e-6, i
        assert tuple(H.to_host(outs[i][1])) == (wx, wy) == tuple(bpk[i]), i
        assert (wx - m // 2, wy - m // 2) == tuple(lags[i]), i


def _spec_host_3d(buf, n):
    a = H.to_host(buf)[: 2 * n[2] * (n[0] // 2 + 1) * n[1]].astype(np.float64)
    return (a[0::2] + 1j * a[1::2]).reshape(n[2], n[0] // 2 + 1, n[1])


def test_half_3d_fwd_256_rows():
    """<float, __half, __half2, 3> forward of an FP16 volume whose 256-point x-pass takes FP16 rows through
    the tile kernel (spectra stored transposed), against the oracle on the same FP16-rounded input."""
    vin, n = (40, 20, 12), (256, 32, 16)
    H.log(f"<float, __half, __half2, 3> fwd {vin} -> {n}")
    rng = np.random.default_rng(41)
    v = rng.standard_normal(vin[::-1]).astype(np.float16)
    ft = H.make_plan(vin, n, n, n, input_type="f16", image_type="c32h", rank=3)
    a = H.to_dev(H.padded_real(v, vin[0], np.float16).reshape(-1))
    out = H.dev_zeros(ft.ReturnFwdOutputMemorySize(), "f16")
    pre = 1.0 / np.sqrt(vin[0] * vin[1] * vin[2])
    ft.FwdFFT(a, out, pre_op=H.F.Scale(pre))
    got = _spec_host_3d(out, n)
    want = O.fwd_3d(H.padded_real(v.astype(np.float32), vin[0]), vin[0], *n) * pre
    assert O.rel_l2(got, want) < TOL_F16


@pytest.mark.parametrize("tin,n,nout,origin", [((8, 8, 8), (32, 32, 32), (16, 16, 16), (-8, -8, -8)),
                                                ((16, 12, 10), (64, 32, 32), (64, 32, 32), (0, 0, 0))])
def test_half_3d_fwd_and_xcorr(tin, n, nout, origin):
    """FourierTransformer<float, __half, __half2, 3>: FP16 volumes, the padded 3-D forward transform stored
    as __half2, and the 3-D correlation of an FP16 template against that __half2 spectrum with an FP16
    trimmed output.  The oracle takes the same FP16-rounded inputs (and, for the correlation, the
    __half2 spectrum the library stored), so the error is that of the FP16 outputs."""
    H.log(f"<float, __half, __half2, 3> fwd and xcorr {tin} in {n} -> {nout} at {origin}")
    rng = np.random.default_rng(sum(n) + sum(tin))
    img = rng.standard_normal(n[::-1]).astype(np.float16)            # [z][y][x]
    tmpl = rng.standard_normal(tin[::-1]).astype(np.float16)
    ntot = n[0] * n[1] * n[2]
    # image spectrum: FP16 volume -> __half2 spectrum, pre-scaled so that FP16 holds it
    mft = H.make_plan(n, n, n, n, input_type="f16", image_type="c32h", rank=3)
    ib = H.to_dev(H.padded_real(img, n[0], np.float16).reshape(-1))
    spec_buf = H.dev_zeros(mft.ReturnFwdOutputMemorySize(), "f16")
    pre = 1.0 / np.sqrt(ntot)
    mft.FwdFFT(ib, spec_buf, pre_op=H.F.Scale(pre))
    got_spec = _spec_host_3d(spec_buf, n)
    want_spec = O.fwd_3d(H.padded_real(img.astype(np.float32), n[0]), n[0], *n) * pre
    err_spec = O.rel_l2(got_spec, want_spec)
    print(f"fwd 3d __half2 spectrum rel_l2={err_spec:.3e}")
    assert err_spec < TOL_F16
    # correlation: FP16 template padded to n, against the stored __half2 spectrum, FP16 output
    ft = H.make_plan(tin, n, n, nout, origin=origin, input_type="f16", image_type="c32h", rank=3)
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()), "f16")
    tp = H.padded_real(tmpl, tin[0], np.float16).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    out = H.dev_poison(ft.ReturnInvOutputMemorySize(), "f16")
    post = 1.0 / (ntot * np.sqrt(tin[0] * tin[1] * tin[2]))
    ft.FwdImageInvFFT(a, spec_buf, out, post_op=H.F.Scale(post))
    got = H.real_window(H.to_host(out).astype(np.float32), *nout)
    spec_t = O.fwd_3d(H.padded_real(tmpl.astype(np.float32), tin[0]), tin[0], *n)
    want = O.inv_3d(got_spec * np.conj(spec_t), *n, *nout, *origin) * post
    err = O.rel_l2(got, want)
    print(f"xcorr 3d FP16 output rel_l2={err:.3e}")
    assert err < TOL_F16
