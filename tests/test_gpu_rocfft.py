"""Second check against a vendor FFT: rocFFT (through torch.fft on ROCm) on the fully padded problem.

The 4096^2 x 384^2 -> 512^2 correlation is recomputed without any pruning: the template zero padded to 4096^2,
R2C, multiplied by conj of itself against the image spectrum, C2R of the full 4096^2 frame, and
the centred 512^2 window cropped.  Both sides run in FP32 on the GPU; the library's spectrum is
also compared with rocFFT's R2C of the micrograph (transposed to the library's C[kx][ky] layout).
"""
import numpy as np
import pytest

from oracle import oracle as O

from . import gpu_helpers as H

pytestmark = pytest.mark.gpu


def test_xcorr_4096_384_against_rocfft():
    from fastfft_amd import workloads as W

    torch, d = H.torch_dev()
    n, t, m = 4096, 384, 512
    H.log("4096^2 x 384^2 -> 512^2 correlation against rocFFT (torch.fft), fully padded")
    mic, tmpl, lags = W.xcorr_config()
    # library: micrograph spectrum C[kx][ky] and the trimmed centred correlation
    mft = H.make_plan((n, n, 1), (n, n, 1), (n, n, 1), (n, n, 1))
    mbuf = H.to_dev(W.to_padded(mic).reshape(-1))
    mft.FwdFFT(mbuf)
    ft = H.make_plan((t, t, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0))
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = W.to_padded(tmpl).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, mbuf, out)
    got = H.real_window(H.to_host(out), m, m)
    # rocFFT: rfft2 over [y][x] gives X[ky][kx]; the library's spectrum is X[kx][ky]
    w = n // 2 + 1
    mic_t = torch.from_numpy(mic).to(d)
    spec_r = torch.fft.rfft2(mic_t)  # [ky][kx], complex64
    spec_lib = torch.view_as_complex(mbuf[: 2 * w * n].view(w, n, 2))  # [kx][ky]
    rel_spec = (torch.linalg.norm(spec_lib - spec_r.T) / torch.linalg.norm(spec_r)).item()
    tpad = torch.zeros((n, n), dtype=torch.float32, device=d)
    tpad[:t, :t] = torch.from_numpy(tmpl).to(d)
    corr = torch.fft.irfft2(spec_r * torch.conj(torch.fft.rfft2(tpad)), s=(n, n)) * (n * n)  # unnormalised
    idx = (torch.arange(m, device=d) - m // 2) % n
    want = corr[idx][:, idx].double().cpu().numpy()
    err = O.rel_l2(got, want)
    print(f"rocFFT check: spectrum rel_l2={rel_spec:.3e} correlation rel_l2={err:.3e}")
    assert rel_spec < 2e-6
    assert err < 2e-6
    gx, gy, _ = O.argmax_2d(got)
    wx, wy, _ = O.argmax_2d(want)
    assert (gx, gy) == (wx, wy) and (gx - m // 2, gy - m // 2) == tuple(lags[0])
