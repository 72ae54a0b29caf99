"""Batched correlation (many templates of uneven sizes against one spectrum) through the C ABI.

Each output of a batch must agree with the single-template FwdImageInvFFT of the same template (the
batched kernels use the sub-transform sizes of the largest template, so agreement is to rounding), each
peak must equal the argmax of its output, and the outputs and peaks of the small batches and of the batches
on a reduced frame are also checked against the FP64 oracle of the whole N-frame pipeline.  The batch is
replayed from a HIP graph with eager work in between.

Where the window and the largest template fit a frame of at most a quarter of the image (Fx >= Lx + Mx - 1,
Fy >= Ly + My - 1), the batch reduces the image to that frame once per call and correlates every template
on it; the "full_frame" mode switches that off, so both forms are compared with the same single
correlations.  The image spectra are random complex numbers (not Hermitian on the DC and Nyquist
columns), so the test also covers how the C2R pass treats those columns in both forms.
"""
import numpy as np
import pytest

from oracle import oracle as O

from . import gpu_helpers as H

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _setup(sizes, n, nout, origin, seed=0):
    from fastfft_amd import workloads as W

    torch, d = H.torch_dev()
    rng = np.random.default_rng(seed)
    mx = max(s[0] for s in sizes), max(s[1] for s in sizes)
    ft = H.make_plan((*mx, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0))
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((w, n[1])) + 1j * rng.standard_normal((w, n[1]))).astype(np.complex64)
    img = H.to_dev(H.complex_to_f32(spec))
    tmpls, hosts, outs = [], [], []
    for i, (x, y) in enumerate(sizes):
        t = W.masked_template(x, 0.45 * min(x, y), seed=seed + 100 + i)[:y, :x] if x == y else \
            rng.standard_normal((y, x)).astype(np.float32)
        hosts.append(t)
        tmpls.append(H.to_dev(H.padded_real(t, x).reshape(-1)))
        outs.append(H.dev_poison(ft.ReturnInvOutputMemorySize()))
    peaks = torch.full((2 * len(sizes),), -1, dtype=torch.int32, device=d)
    vals = torch.zeros(len(sizes), dtype=torch.float32, device=d)
    ft.SetBatch(tmpls, sizes, outs, peaks, vals)
    return torch, ft, spec, img, tmpls, hosts, outs, peaks, vals


BATCHES = [([(128, 128), (256, 256), (384, 384), (128, 128), (384, 384)], (4096, 4096), (512, 512), (-256, -256)),
           ([(200, 200), (160, 120), (97, 200)], (2048, 2048), (150, 200), (-70, 30)),
           ([(40, 30), (64, 64), (17, 50)], (256, 128), (64, 32), (-32, -16)),
           ([(100, 60), (100, 60)], (256, 128), (100, 60), (-50, -30)),
           # 1024-wide frames of 2048- and 8192-point images: the frame's fused C2R + R2C rows at Q = 2 and 8
           ([(256, 256), (200, 180)], (2048, 2048), (512, 400), (-256, -200)),
           ([(300, 200), (128, 128)], (8192, 1024), (512, 64), (-256, -32))]

# frame of the batches that run on a reduced frame
FRAMES = {((4096, 4096), (512, 512)): "1024x1024", ((2048, 2048), (150, 200)): "512x512",
          ((2048, 2048), (512, 400)): "1024x1024", ((8192, 1024), (512, 64)): "1024x512"}


@pytest.mark.parametrize("mode", ["tuned", "chunk2", "no_packed", "no_frame_fusion", "row_major", "frame_rows1",
                                  "frame_rows2", "full_frame", "per_template"])
@pytest.mark.parametrize("sizes,n,nout,origin", BATCHES)
def test_batch_matches_single_correlations(sizes, n, nout, origin, mode, monkeypatch):
    """Batched kernels (chunk size from the tuner, or 2 to force several chunks, or the two-wave row
    kernels in place of the packed ones, or the frame's C2R and R2C rows as two kernels) and the
    per-template launches all agree with single-template correlations and give the argmax of their own
    outputs."""
    if mode == "no_frame_fusion":
        monkeypatch.setenv("FASTFFT_NO_FRAME_FUSION", "1")
    if mode == "chunk2":
        monkeypatch.setenv("FASTFFT_BATCH_CHUNK", "2")
    if mode == "no_packed":
        monkeypatch.setenv("FASTFFT_NO_PACKED_ROWS", "1")
    if mode == "per_template":
        monkeypatch.setenv("FASTFFT_NO_WAVE_ROWS", "1")
    if mode == "full_frame":
        monkeypatch.setenv("FASTFFT_NO_SUBFRAME", "1")
    if mode == "row_major":
        monkeypatch.setenv("FASTFFT_BATCH_ROW_MAJOR", "1")
    if mode.startswith("frame_rows"):
        monkeypatch.setenv("FASTFFT_FRAME_ROWS", mode[-1])
    H.log(f"batch {sizes} in {n} -> {nout} at {origin} mode={mode}")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=len(sizes))
    form = ft.Describe("batch").splitlines()[0]
    framed = ("tuned", "chunk2", "no_packed", "no_frame_fusion", "row_major", "frame_rows1", "frame_rows2")
    frame = FRAMES.get((n, nout)) if mode in framed else None
    assert f"frame={frame or 'full'}" in form, form
    if frame:
        # the frame's template and output slots: column-major where the tiled row kernels cover the
        # frame's rows (512- and 1024-point rows), row-major when asked for
        cm = mode != "row_major" and frame.split("x")[0] in ("512", "1024") and mode != "no_packed"
        assert ("slots=column-major" if cm else "slots=row-major") in form, form
    ft.FwdImageInvFFTBatch(img)
    pk, vv = H.to_host(peaks).reshape(-1, 2), H.to_host(vals)
    pitch = H.F.padded_pitch(nout[0])
    for i, (x, y) in enumerate(sizes):
        single = H.make_plan((x, y, 1), (*n, 1), (*n, 1), (*nout, 1), origin=(*origin, 0))
        a = H.dev_zeros(max(single.ReturnInputMemorySize(), single.ReturnInvOutputMemorySize()))
        a[: tmpls[i].numel()] = tmpls[i]
        o = H.dev_zeros(single.ReturnInvOutputMemorySize())
        single.FwdImageInvFFT(a, img, o)
        got = H.real_window(H.to_host(outs[i]), *nout)
        assert O.rel_l2(got, H.real_window(H.to_host(o), *nout)) < TOL, i
        rx, ry, rv = single.ArgMax(outs[i], nout[0], nout[1], pitch)  # argmax of the batch's own output
        assert tuple(pk[i]) == (rx, ry) and vv[i] == np.float32(rv), i
        # the FP64 oracle of the full N-frame pipeline: every small batch, and the batches on a reduced
        # frame in the modes that run the frame (the fused frame rows at Q = 2, 4 and 8 included)
        if n[0] * n[1] <= 256 * 128 or (frame and mode != "chunk2"):
            want = O.xcorr_2d(H.padded_real(hosts[i], x), x, spec, n[0], n[1], nout[0], nout[1], *origin)
            assert O.rel_l2(got, want) < TOL, i
            wx, wy, _ = O.argmax_2d(want)
            assert tuple(pk[i]) == (wx, wy), i


@pytest.mark.parametrize("rows", [1, 2, 4])
@pytest.mark.parametrize("case", [0, 4, 5])
def test_frame_fusion_matches_two_kernels(case, rows, monkeypatch):
    """The frame's fused C2R + R2C rows (Q = 4, 2 and 8 for the 4096-, 2048- and 8192-point images; one, two
    or four rows per workgroup, at most two at Q = 8) give the outputs and peaks of the same batch with the
    two kernels run one after the other, to FP32 rounding."""
    sizes, n, nout, origin = BATCHES[case]
    H.log(f"frame fusion against two kernels, {sizes} in {n} -> {nout}, {rows} rows per workgroup")
    monkeypatch.setenv("FASTFFT_FRAME_ROWS", str(rows))
    res = {}
    for fused in (True, False):
        if not fused:
            monkeypatch.setenv("FASTFFT_NO_FRAME_FUSION", "1")
        torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=17)
        assert f"frame={FRAMES[(n, nout)]}" in ft.Describe("batch")
        ft.FwdImageInvFFTBatch(img)
        res[fused] = ([H.real_window(H.to_host(o), *nout) for o in outs], H.to_host(peaks).copy())
    for a, b in zip(res[True][0], res[False][0]):
        assert O.rel_l2(a, b) < 1e-6
    assert np.array_equal(res[True][1], res[False][1])


@pytest.mark.parametrize("case", [1, 2])
def test_batch_graph_replays(case):
    """Replays of a captured batch equal the eager batch, also for a batch on a reduced frame."""
    sizes, n, nout, origin = BATCHES[case]
    H.log(f"batch {sizes} in {n} -> {nout} in a HIP graph, replayed with eager work in between")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=9)
    side = torch.cuda.Stream()
    ft.SetStream(side.cuda_stream)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        ft.FwdImageInvFFTBatch(img)
    torch.cuda.synchronize()
    want = [H.to_host(o).copy() for o in outs]
    want_pk, want_v = H.to_host(peaks).copy(), H.to_host(vals).copy()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        ft.FwdImageInvFFTBatch(img)
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        peaks.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, o in enumerate(outs):
            assert np.array_equal(H.to_host(o), want[i], equal_nan=True), (rep, i)
        assert np.array_equal(H.to_host(peaks), want_pk) and np.array_equal(H.to_host(vals), want_v), rep
        ft.FwdImageInvFFTBatch(img)
        torch.cuda.synchronize()


def test_batch_rejects_oversized_template():
    torch, d = H.torch_dev()
    ft = H.make_plan((64, 64, 1), (256, 256, 1), (256, 256, 1), (64, 64, 1))
    t = H.dev_zeros(H.F.padded_pitch(128) * 128)
    o = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    with pytest.raises(H.F.FastFFTError) as e:
        ft.SetBatch([t], [(128, 128)], [o])
    assert e.value.status == H.F.INVALID_ARGUMENT


@pytest.mark.parametrize("case", [0, 1])
def test_batch_parts_in_order_equal_the_call(case, monkeypatch):
    """fastfft_run_batch_stage 0..4 in order (frame, R2C rows, correlation pass, C2R rows, argmax) gives
    the outputs and peaks of one FwdImageInvFFTBatch call, bit for bit, for a batch of one chunk."""
    monkeypatch.setenv("FASTFFT_BATCH_CHUNK", "8")
    sizes, n, nout, origin = BATCHES[case]
    H.log(f"batch parts in order, {sizes} in {n} -> {nout}")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=13)
    assert "chunk=8" in ft.Describe("batch")
    ft.FwdImageInvFFTBatch(img)
    torch.cuda.synchronize()
    want = [H.to_host(o).copy() for o in outs]
    want_pk = H.to_host(peaks).copy()
    for o in outs:
        o.fill_(float("nan"))
    peaks.fill_(-1)
    for part in range(5):
        ft.RunBatchStage(part, img)
    torch.cuda.synchronize()
    for i, o in enumerate(outs):
        assert np.array_equal(H.to_host(o), want[i], equal_nan=True), i
    assert np.array_equal(H.to_host(peaks), want_pk)


def test_batch_256_uneven_templates_4096():
    """The full uneven batch: 96 templates of 128^2, 96 of 256^2 and 64 of 384^2 (circular masks) against one
    4096^2 spectrum, 512^2 centred window.  Every peak equals the peak of the template's single correlation,
    and every 16th output equals its single correlation (rel-L2 within 2e-6)."""
    from fastfft_amd import workloads as W

    H.log("batch of 256 uneven templates against 4096^2")
    torch, d = H.torch_dev()
    n, m = 4096, 512
    mic, _, _ = W.xcorr_config()
    mft = H.make_plan((n, n, 1), (n, n, 1), (n, n, 1), (n, n, 1))
    img = H.dev_zeros(mft.ReturnFwdOutputMemorySize())
    img[:] = H.to_dev(W.to_padded(mic).reshape(-1))
    mft.FwdFFT(img)
    tmpls = W.batched_templates()
    sizes = [(t.shape[1], t.shape[0]) for t in tmpls]
    ft = H.make_plan((384, 384, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0))
    tb = [H.to_dev(W.to_padded(t).reshape(-1)) for t in tmpls]
    outs = [H.dev_poison(ft.ReturnInvOutputMemorySize()) for _ in tmpls]
    peaks = torch.full((2 * len(tmpls),), -1, dtype=torch.int32, device=d)
    vals = torch.zeros(len(tmpls), dtype=torch.float32, device=d)
    ft.SetBatch(tb, sizes, outs, peaks, vals)
    assert "frame=1024x1024" in ft.Describe("batch")
    ft.FwdImageInvFFTBatch(img)
    pk = H.to_host(peaks).reshape(-1, 2)
    singles = {}
    for i, (x, y) in enumerate(sizes):
        if (x, y) not in singles:
            singles[(x, y)] = H.make_plan((x, y, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0))
        s = singles[(x, y)]
        a = H.dev_zeros(max(s.ReturnInputMemorySize(), s.ReturnInvOutputMemorySize()))
        a[: tb[i].numel()] = tb[i]
        o = H.dev_zeros(s.ReturnInvOutputMemorySize())
        sp = torch.zeros(2, dtype=torch.int32, device=d)
        sv = torch.zeros(1, dtype=torch.float32, device=d)
        s.FwdImageInvFFTArgMax(a, img, o, sp, sv)
        assert tuple(pk[i]) == tuple(H.to_host(sp)), i
        if i % 16 == 0:
            got = H.real_window(H.to_host(outs[i]), m, m)
            assert O.rel_l2(got, H.real_window(H.to_host(o), m, m)) < TOL, i
    # one template of each size and the last one against the FP64 oracle of the 4096^2 pipeline
    mspec = H.complex_from_f32(H.to_host(img), (n // 2 + 1, n))
    for i in (0, 96, 192, len(tmpls) - 1):
        x, y = sizes[i]
        got = H.real_window(H.to_host(outs[i]), m, m)
        want = O.xcorr_2d(W.to_padded(tmpls[i]), x, mspec, n, n, m, m, -m // 2, -m // 2)
        assert O.rel_l2(got, want) < TOL, i
        wx, wy, _ = O.argmax_2d(want)
        assert tuple(pk[i]) == (wx, wy), i


def _planted_mixed(n, m, sizes, seed):
    """Micrograph noise plus 0.5 x template i at a seeded lag inside the centred m x m window."""
    from fastfft_amd import workloads as W

    rng = np.random.default_rng(seed)
    tmpls = [W.masked_template(t, 0.42 * t, seed=seed + 1 + i) for i, t in enumerate(sizes)]
    mic = rng.standard_normal((n, n)).astype(np.float32)
    lags = []
    for tp in tmpls:
        t = tp.shape[0]
        lx, ly = int(rng.integers(-m // 2 + 8, m // 2 - 8)), int(rng.integers(-m // 2 + 8, m // 2 - 8))
        lags.append((lx, ly))
        ys, xs = (np.arange(t) + ly) % n, (np.arange(t) + lx) % n
        mic[np.ix_(ys, xs)] += 0.5 * tp
    return mic, tmpls, lags


@pytest.mark.parametrize("chunk", [0, 4, "row_major"])
def test_headline_batch_4096_mixed_against_oracle(chunk, monkeypatch):
    """The headline form: FP32 templates of 128^2, 256^2 and 384^2 (two of each, circular masks, each planted
    once in a 4096^2 micrograph at a lag inside the window) correlated by one batched call on the reduced
    1024^2 frame, 512^2 window of lags -256..255.  Every output is within 2e-6 (relative L2) of the FP64
    oracle of the full 4096^2 pipeline, and every peak equals the oracle's and the planted lag.  chunk 4
    runs the batch in two chunks of launches."""
    from fastfft_amd import workloads as W

    if chunk == "row_major":
        monkeypatch.setenv("FASTFFT_BATCH_ROW_MAJOR", "1")
        chunk = 0
    elif chunk:
        monkeypatch.setenv("FASTFFT_BATCH_CHUNK", str(chunk))
    n, m = 4096, 512
    H.log(f"headline batch 4096^2 x (128, 256, 384)^2 -> 512^2 against the FP64 oracle, chunk={chunk or 'tuned'}")
    torch, d = H.torch_dev()
    tsz = [128, 256, 384, 128, 256, 384]
    mic, tmpls, lags = _planted_mixed(n, m, tsz, seed=71)
    mft = H.make_plan((n, n, 1), (n, n, 1), (n, n, 1), (n, n, 1))
    img = H.dev_zeros(mft.ReturnFwdOutputMemorySize())
    img[:] = H.to_dev(W.to_padded(mic).reshape(-1))
    mft.FwdFFT(img)
    mspec = H.complex_from_f32(H.to_host(img), (n // 2 + 1, n))
    ft = H.make_plan((384, 384, 1), (n, n, 1), (n, n, 1), (m, m, 1), origin=(-m // 2, -m // 2, 0))
    tb = [H.to_dev(W.to_padded(t).reshape(-1)) for t in tmpls]
    outs = [H.dev_poison(ft.ReturnInvOutputMemorySize()) for _ in tmpls]
    peaks = torch.full((2 * len(tmpls),), -1, dtype=torch.int32, device=d)
    vals = torch.zeros(len(tmpls), dtype=torch.float32, device=d)
    ft.SetBatch(tb, [(t, t) for t in tsz], outs, peaks, vals)
    form = ft.Describe("batch").splitlines()[0]
    assert "frame=1024x1024" in form and "precision=fp32" in form, form
    if chunk:
        assert f"chunk={chunk}" in form, form
    ft.FwdImageInvFFTBatch(img)
    pk, vv = H.to_host(peaks).reshape(-1, 2), H.to_host(vals)
    for i, tp in enumerate(tmpls):
        t = tsz[i]
        got = H.real_window(H.to_host(outs[i]), m, m)
        want = O.xcorr_2d(W.to_padded(tp), t, mspec, n, n, m, m, -m // 2, -m // 2)
        err = O.rel_l2(got, want)
        print(f"template {i} ({t}^2): rel_l2={err:.3e} max_abs={O.max_abs(got, want):.3e}")
        assert err < TOL, i
        wx, wy, wv = O.argmax_2d(want)
        assert tuple(pk[i]) == (wx, wy), i
        assert (wx - m // 2, wy - m // 2) == lags[i], i
        assert abs(float(vv[i]) - wv) <= 2e-6 * abs(wv), i
