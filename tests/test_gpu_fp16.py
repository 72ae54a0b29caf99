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
