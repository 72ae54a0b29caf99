"""Template-matching statistics of the batched correlation (fastfft_set_batch_stats).

With statistics set, a batched call updates per-pixel maps in place of writing the correlation maps: the
maximum over templates, the template id of its first occurrence, the sum and the sum of squares.  The
maps are checked against the FP64 oracle's correlation maps and against the same batch's own maps (the
batch run without statistics), both reduced on the host in template order in FP64: maximum and sums within FP32 rounding, the template id
exactly wherever the two largest values are further apart than that rounding.  Also covered:
accumulation over two calls and over several chunks, graph replays, the reset, untouched outputs, and the
batches the statistics form does not take.
"""
import numpy as np
import pytest

from . import gpu_helpers as H
from .test_gpu_batch import BATCHES, _setup

pytestmark = pytest.mark.gpu

# a 1024-point and a 512-point frame along x after the frame reduction
CASES = [0, 1]


def _maps(ft, img, outs, nout):
    ft.FwdImageInvFFTBatch(img)
    return np.stack([H.real_window(H.to_host(o), nout[0], nout[1]) for o in outs])  # [t][y][x], FP64


def _stats_tensors(torch, nout):
    d = torch.device("cuda:0")
    n = nout[0] * nout[1]
    return (torch.zeros(n, dtype=torch.float32, device=d), torch.zeros(n, dtype=torch.int32, device=d),
            torch.zeros(n, dtype=torch.float32, device=d), torch.zeros(n, dtype=torch.float32, device=d))


def _check(maps, first_ids, mip, best, s, sq, nout):
    """maps: list over calls of [t][y][x]; first_ids: the id of template 0 in each call."""
    allm = np.concatenate(maps)
    ids = np.concatenate([fid + np.arange(m.shape[0]) for fid, m in zip(first_ids, maps)])
    scale = np.abs(allm).max()
    mip = H.to_host(mip).reshape(nout[1], nout[0])
    best = H.to_host(best).reshape(nout[1], nout[0])
    s = H.to_host(s).reshape(nout[1], nout[0])
    sq = H.to_host(sq).reshape(nout[1], nout[0])
    want_mip = allm.max(axis=0)
    assert np.abs(mip - want_mip).max() <= 2e-6 * scale
    # pixels with a clear winner among the distinct templates (later calls repeat the first call's maps)
    top = np.sort(maps[0], axis=0)
    clear = (top[-1] - top[-2]) > 1e-5 * scale
    want_best = ids[np.argmax(allm, axis=0)]  # first maximum in template order
    assert clear.mean() > 0.9
    assert np.array_equal(best[clear], want_best[clear])
    assert np.abs(s - allm.sum(axis=0)).max() <= 1e-5 * np.abs(allm).sum(axis=0).max()
    assert np.abs(sq - (allm ** 2).sum(axis=0)).max() <= 1e-5 * (allm ** 2).sum(axis=0).max()


@pytest.mark.parametrize("mode", ["tuned", "chunk2", "groups"])
@pytest.mark.parametrize("case", CASES)
def test_batch_stats_match_the_maps(case, mode, monkeypatch):
    """chunk2: several chunks per call; groups: one template per group (partial maps merged in order)."""
    if mode == "chunk2":
        monkeypatch.setenv("FASTFFT_BATCH_CHUNK", "2")
    if mode == "groups":
        monkeypatch.setenv("FASTFFT_STATS_PER_GROUP", "1")
    sizes, n, nout, origin = BATCHES[case]
    H.log(f"batch stats {sizes} in {n} -> {nout} at {origin} mode={mode}")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=21 + case)
    maps = _maps(ft, img, outs, nout)
    mip, best, s, sq = _stats_tensors(torch, nout)
    ft.SetBatchStats(mip, best, s, sq, first_id=10)
    assert "stats=on first_id=10" in ft.Describe("batch").splitlines()[0]
    ft.ResetBatchStats()
    for o in outs:
        o.fill_(7.0)
    ft.FwdImageInvFFTBatch(img)
    _check([maps], [10], mip, best, s, sq, nout)
    # the outputs are not written in this form
    assert all(np.all(H.to_host(o) == 7.0) for o in outs)
    # a second call accumulates, with the ids of its own templates; ties keep the first call
    ft.SetBatchStats(mip, best, s, sq, first_id=10 + len(sizes))
    ft.FwdImageInvFFTBatch(img)
    _check([maps, maps], [10, 10 + len(sizes)], mip, best, s, sq, nout)
    # switched off, the batch writes its maps and peaks again
    ft.SetBatchStats()
    assert "stats=on" not in ft.Describe("batch")
    again = _maps(ft, img, outs, nout)
    assert np.array_equal(again, maps)


@pytest.mark.parametrize("case", CASES)
def test_batch_stats_against_oracle_maps(case):
    """The statistics of one call against the FP64 oracle: every template's correlation map from the
    oracle of the full N-frame pipeline, reduced on the host in FP64 (maximum, first template reaching it,
    sum, sum of squares).  Case 0 is the 4096^2 spectrum on the reduced 1024^2 frame with 128^2, 256^2
    and 384^2 templates and the centred 512^2 window."""
    from oracle import oracle as O

    sizes, n, nout, origin = BATCHES[case]
    H.log(f"batch stats {sizes} in {n} -> {nout} at {origin} against oracle maps")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=51 + case)
    mip, best, s, sq = _stats_tensors(torch, nout)
    ft.SetBatchStats(mip, best, s, sq, first_id=3)
    ft.ResetBatchStats()
    ft.FwdImageInvFFTBatch(img)
    want = np.stack([O.xcorr_2d(H.padded_real(t, x), x, spec, n[0], n[1], nout[0], nout[1], *origin)
                     for t, (x, y) in zip(hosts, sizes)])  # [t][y][x], FP64
    _check([want], [3], mip, best, s, sq, nout)


@pytest.mark.parametrize("case", CASES)
def test_batch_stats_graph_replays(case):
    """A captured reset + batch, replayed with eager work in between, equals the eager result."""
    sizes, n, nout, origin = BATCHES[case]
    H.log(f"batch stats {sizes} in {n} -> {nout} in a HIP graph")
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=31 + case)
    mip, best, s, sq = _stats_tensors(torch, nout)
    ft.SetBatchStats(mip, best, s, sq, first_id=0)
    side = torch.cuda.Stream()
    ft.SetStream(side.cuda_stream)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        ft.ResetBatchStats()
        ft.FwdImageInvFFTBatch(img)
    torch.cuda.synchronize()
    want = [H.to_host(a).copy() for a in (mip, best, s, sq)]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        ft.ResetBatchStats()
        ft.FwdImageInvFFTBatch(img)
    for rep in range(3):
        for a in (mip, best, s, sq):
            a.fill_(3)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, w in zip((mip, best, s, sq), want):
            assert np.array_equal(H.to_host(a), w), rep
        with torch.cuda.stream(side):
            ft.ResetBatchStats()
            ft.FwdImageInvFFTBatch(img)
        torch.cuda.synchronize()


def test_batch_stats_rejects_uncovered_batches():
    """A batch on 256-point rows (no statistics form) and partial pointer sets are refused."""
    import fastfft_amd as F

    sizes, n, nout, origin = BATCHES[2]
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=3)
    mip, best, s, sq = _stats_tensors(torch, nout)
    with pytest.raises(F.FastFFTError):
        ft.SetBatchStats(mip, best, s, sq)
    sizes, n, nout, origin = BATCHES[1]
    torch, ft, spec, img, tmpls, hosts, outs, peaks, vals = _setup(sizes, n, nout, origin, seed=4)
    mip, best, s, sq = _stats_tensors(torch, nout)
    with pytest.raises(F.FastFFTError):
        ft.SetBatchStats(mip, best, None, sq)
    with pytest.raises(F.FastFFTError):
        ft.ResetBatchStats()
