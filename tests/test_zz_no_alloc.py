"""Allocation-free correlation entry point (fastfft_fwd_image_inv_argmax_noalloc).

Its intermediates and argmax slots live in the library's static device scratch; the kernels are those of
fastfft_fwd_image_inv_argmax, so its window, peak and value must equal that entry point's bit for bit.
Two calls in a row, the second on a smaller plan (the scratch keeps the first call's contents), then the
same two plans through the allocating entry point.  The first call of the process may be the
allocation-free one: it also fills the device's twiddle table.
"""
import numpy as np
import pytest

from . import gpu_helpers as H

pytestmark = pytest.mark.gpu


def _case(n, t, m, origin, seed):
    """A random half spectrum of an n x n image, a t = (x, y) template, an m = (x, y) window at origin."""
    rng = np.random.default_rng(seed)
    spec = (rng.standard_normal((n // 2 + 1, n)) + 1j * rng.standard_normal((n // 2 + 1, n))).astype(np.complex64)
    tmpl = rng.standard_normal((t[1], t[0])).astype(np.float32)
    ft = H.make_plan((t[0], t[1], 1), (n, n, 1), (n, n, 1), (m[0], m[1], 1), origin=(origin[0], origin[1], 0))
    img = H.to_dev(H.complex_to_f32(spec))
    tp = H.padded_real(tmpl, t[0]).reshape(-1)
    return ft, img, tp


def _run(ft, img, tp, noalloc):
    torch, d = H.torch_dev()
    tbuf = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tbuf[: tp.size] = H.to_dev(tp)
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    peak = torch.full((2,), -7, dtype=torch.int32, device=d)
    val = torch.full((1,), float("nan"), dtype=torch.float32, device=d)
    fn = ft.FwdImageInvFFTArgMaxNoAlloc if noalloc else ft.FwdImageInvFFTArgMax
    fn(tbuf, img, out, peak, val)
    return H.to_host(out).copy(), H.to_host(peak).copy(), H.to_host(val).copy()


CASES = [
    # the single-correlation workload: 384^2 template, 4096^2 spectrum, 512^2 window centred on lag 0
    dict(n=4096, t=(384, 384), m=(512, 512), origin=(-256, -256), seed=71),
    # a smaller plan next: odd template, rectangular window off centre
    dict(n=1024, t=(100, 90), m=(128, 96), origin=(-40, -30), seed=72),
]


def test_noalloc_matches_allocating_entry_point():
    cases = [_case(**c) for c in CASES]
    got = []
    for c, (ft, img, tp) in zip(CASES, cases):
        H.log(f"noalloc correlation n={c['n']} t={c['t']} m={c['m']} origin={c['origin']}")
        got.append(_run(ft, img, tp, noalloc=True))
    for i, (c, (ft, img, tp)) in enumerate(zip(CASES, cases)):
        H.log(f"allocating correlation n={c['n']} t={c['t']} m={c['m']} origin={c['origin']}")
        want = _run(ft, img, tp, noalloc=False)
        mx, my = c["m"]
        win = lambda a: H.real_window(a, mx, my)
        assert np.isfinite(win(got[i][0])).all(), i
        assert np.array_equal(win(got[i][0]), win(want[0])), i
        assert np.array_equal(got[i][1], want[1]) and np.array_equal(got[i][2], want[2]), i
        # the peak is the window's first maximum
        w = win(want[0])
        y, x = np.unravel_index(int(np.argmax(w)), w.shape)
        assert (int(want[1][0]), int(want[1][1])) == (x, y), i


def test_noalloc_repeat_and_other_stream():
    """Three calls on one plan give identical results (nothing of the scratch needs to start at zero),
    also on a stream of the caller's."""
    torch, _ = H.torch_dev()
    c = CASES[1]
    ft, img, tp = _case(**c)
    first = _run(ft, img, tp, noalloc=True)
    again = _run(ft, img, tp, noalloc=True)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        third = _run(ft, img, tp, noalloc=True)
    for a, b in ((first, again), (first, third)):
        assert all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))
