"""The CPU baseline bench.py times (oracle/cpu_baseline.c through the FFTW3 API, or scipy.fft) computes
the same windowed correlation as the FP64 oracle: relative L2 at FP32 level and the same peak."""
import numpy as np

from oracle import cpu_baseline as C
from oracle import oracle as O


def _to_padded(a):
    out = np.zeros((a.shape[0], 2 * (a.shape[1] // 2 + 1)), np.float32)
    out[:, : a.shape[1]] = a
    return out


def test_cpu_baseline_matches_oracle():
    n, t, m = 256, 40, 64
    rng = np.random.default_rng(5)
    mic = rng.standard_normal((n, n)).astype(np.float32)
    tmpl = rng.standard_normal((t, t)).astype(np.float32)
    ys, xs = (np.arange(t) + 10) % n, (np.arange(t) - 6) % n
    mic[np.ix_(ys, xs)] += 0.8 * tmpl  # a peak at lag (-6, 10) inside the window
    r = C.xcorr(mic, tmpl, m, nthreads=2, warmup=1, min_corr=2, max_seconds=0.0)
    spec = O.fwd_2d(_to_padded(mic), n, n, n)
    want = O.xcorr_2d(_to_padded(tmpl), t, spec, n, n, m, m, -m // 2, -m // 2)
    assert O.rel_l2(r["window"].astype(np.float64), want) < 2e-6, r["library"]
    wx, wy, _ = O.argmax_2d(want)
    assert tuple(r["peak_xy"]) == (wx, wy) == (m // 2 - 6, m // 2 + 10)
    assert r["n_done"] == 2 and r["threads"] == 2 and r["library"]
