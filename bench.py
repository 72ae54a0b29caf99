#!/usr/bin/env python3
"""Benchmark: batched padded + trimmed cross-correlation on one MI355X.

Workload: a 4096 x 4096 FP32 micrograph (N(0,1) noise + 20 copies of the first
template at amplitude 0.5) whose spectrum C[kx][ky] is computed once by FwdFFT
(timed separately, not part of a step); B = 64 FP32 templates of 384 x 384
(N(0,1), radius-160 circular masks; template 0 is the planted one) zero padded
to 4096 x 4096; every output trimmed to the 512 x 512 window of lags
-256..255 centred on lag 0, with its argmax peak.

One step = one FwdImageInvFFTBatch call through the C ABI: all B correlations
(including the per-call reduction of the image to the 1024 x 1024 frame that
the window and templates need) and their peaks, on the GPU with inputs already
resident in HBM, captured once into a HIP graph that is replayed.  value =
B correlations x steps / time.

Beside the headline line the JSON holds:
  split_bf16  the same batch with the correlation pass on the matrix cores
              (fastfft_set_precision), its speedup over FP32 and its parity;
  uneven_batch 256 templates of mixed sizes (96 x 128^2, 96 x 256^2,
              64 x 384^2) in one call per step;
  batch_256   the headline's 384^2 templates, 256 per call (one frame
              reduction for four times as many correlations);
  config5_fp16 256 FP16 512^2 templates against the FP16 micrograph's
              __half2 spectrum with FP16 outputs, FP32 and split-BF16;
  stats_2dtm  the headline batch in template-matching statistics form
              (per-pixel maximum, its template, sum and sum of squares
              updated in place of the maps), against the batch's own maps;
  config4_3d  the 3-D size changes: padded forward 64^3 -> 256^3 and trimmed
              inverse 512^3 -> 128^3 (centred), with their HBM fractions;
  single      one correlation at a time (fastfft_fwd_image_inv_argmax, the
              FwdImageInvFFT pipeline on the full 4096 x 4096 frame);
  roofline    the dominant kernel of a step (the batched correlation pass):
              its average launch duration timed with HIP events around graph
              replays of that part alone (fastfft_run_batch_stage), on the
              stream the graph runs on; traffic from the rocprofv3 counter
              summary committed under profiles/;
  cpu_baseline the same correlation on this host's cores through the FFTW3
              single-precision API (FFTW if installed, else Intel MKL's FFTW3
              interface, else scipy.fft; oracle/cpu_baseline.c): micrograph
              r2c once, then per template pad, r2c, image * conj, c2r, crop,
              argmax, on a bounded sample of template 0's correlations; the
              library, its version, the thread count and the CPU model are
              reported.  The FP64 oracle (oracle/) gives the parity figures.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, T, M = 4096, 384, 512
HBM_PEAK_GBS = 8000.0
# the split-BF16 correlation pass of the 1024-point frame: per column a Toeplitz product of 512 output
# rows by K = 384 template rows for every template, as 3 real products (Gauss's form of the complex
# product) x 3 split terms (hi*hi, hi*lo, lo*hi); v_mfma_f32_32x32x16_bf16 instructions per template column
MFMA_KERNEL = "k_xcorr_toeplitz"
MFMA_PER_COLUMN = 3 * 3 * (M // 32) * (T // 16) / 32
TRAFFIC_PROFILE = "profiles/xcorr_col_frame1024_batch_hbm_traffic.md"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps (batched calls)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=64, help="templates per step")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="bounded CPU-baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph-batch", type=int, default=10, help="correlations per captured graph (single form)")
    ap.add_argument("--only", default="all",
                    help="comma-separated sections besides the headline batch (stats, split_bf16, uneven, config5, "
                         "config4, single, cpu), e.g. for a short profiled run; default all")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library has no multi-GPU path (asked for --gpus {args.gpus}). "
              "A batch would shard by template, one process per GPU; see the design notes.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import fastfft_amd as F
    from fastfft_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    B, K = args.batch, args.steps
    w = N // 2 + 1  # spectrum columns of the micrograph
    only = {x.strip() for x in args.only.split(",") if x.strip()}

    def want(section):
        return "all" in only or section in only

    mic, tmpl, lags = W.xcorr_config(n=N, tsize=T, radius=160.0, out=M)

    # micrograph spectrum, once (reported separately)
    mft = F.FourierTransformer()
    mft.SetForwardFFTPlan(N, N, 1, N, N, 1)
    mft.SetInverseFFTPlan(N, N, 1, N, N, 1)
    spec = torch.from_numpy(W.to_padded(mic).reshape(-1)).to(dev)
    mft.FwdFFT(spec)  # warm
    # ten transforms of freshly copied micrographs, enqueued back to back (GPU time per transform)
    mics = [torch.from_numpy(W.to_padded(mic).reshape(-1)).to(dev) for _ in range(10)]
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for b in mics:
        mft.FwdFFT(b)
    e1.record()
    torch.cuda.synchronize()
    mic_fwd_ms = e0.elapsed_time(e1) / len(mics)
    spec.copy_(mics[0])
    del mics
    # padded real rows in, the (N/2 + 1) x N spectrum out, in place
    mic_fwd_bytes = 4 * N * 2 * (N // 2 + 1) * 2

    conj = F.ConjMul()
    cap = torch.cuda.Stream()

    def capture(fn, reps):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=cap):
            for _ in range(reps):
                fn()
        return g

    def time_graph(g, replays, warm=2):
        for _ in range(warm):
            g.replay()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(cap)
        with torch.cuda.stream(cap):
            for _ in range(replays):
                g.replay()
        b.record(cap)
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    def plan():
        p = F.FourierTransformer()
        p.SetForwardFFTPlan(T, T, 1, N, N, 1)
        p.SetInverseFFTPlan(N, N, 1, M, M, 1)
        p.SetOutputWindow(-M // 2, -M // 2)
        return p

    tbufs = [torch.from_numpy(W.to_padded(tmpl).reshape(-1)).to(dev)]
    tbufs += [torch.from_numpy(W.to_padded(W.masked_template(T, 160.0, seed=1000 + i)).reshape(-1)).to(dev)
              for i in range(1, B)]

    def batch_form(precision):
        bft = plan()
        outs = [torch.zeros(bft.ReturnInvOutputMemorySize(), dtype=torch.float32, device=dev) for _ in range(B)]
        pk = torch.zeros(2 * B, dtype=torch.int32, device=dev)
        val = torch.zeros(B, dtype=torch.float32, device=dev)
        bft.SetPrecision(precision)
        t_set = time.perf_counter()
        bft.SetBatch(tbufs, [(T, T)] * B, outs, pk, val)  # synchronous: picks the chunk size by timing
        set_s = time.perf_counter() - t_set
        bft.SetStream(cap)
        g = capture(lambda: bft.FwdImageInvFFTBatch(spec, intra_op=conj), 1)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        ms = time_graph(g, K, warm=0) / K
        return bft, outs, pk, ms, set_s

    # ---------------------------------------------------------------- headline: batched FP32
    bft, outs, bpk, ms_step, set_s = batch_form("fp32")
    # outputs and peaks of the last replay (the part timings below overwrite the batch's buffers)
    b_out0 = outs[0].cpu().numpy()
    b_pk_all = bpk.cpu().numpy()
    b_pk = [int(v) for v in b_pk_all[:2]]
    head = bft.Describe("batch").splitlines()[0]
    chunk = int(re.search(r"chunk=(\d+)", head).group(1))
    frame = re.search(r"frame=(\d+)x(\d+)", head)
    fx, fy = (int(frame.group(1)), int(frame.group(2))) if frame else (N, N)
    n_launch = (B + chunk - 1) // chunk
    # per-part time of one step: graphs of one part over the whole batch
    part_ms = {}
    reps = max(2, K // 4)
    for i, name in enumerate(["frame", "r2c_rows", "correlation", "c2r_rows", "argmax"]):
        gp = capture(lambda i=i: bft.RunBatchStage(i, spec, intra_op=conj), 1)
        part_ms[name] = time_graph(gp, reps) / reps
    xc_ms = part_ms["correlation"] / n_launch  # average launch of the correlation pass
    # correlation-pass bytes per launch: the frame spectrum's columns once, each template's
    # A[L][w] in and B[M][w] out (complex float)
    wf = fx // 2 + 1
    xc_bytes = 8 * wf * fy + chunk * 8 * wf * (T + M)
    achieved = xc_bytes / (xc_ms * 1e-3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, TRAFFIC_PROFILE)
    if os.path.exists(prof) and (fx, fy) == (1024, 1024):
        for line in open(prof):
            m = re.match(r"traffic_bytes_per_dispatch chunk=(\d+): (\d+)", line)
            if m and int(m.group(1)) == chunk:
                traffic = int(m.group(2))
    bft.SetStream(torch.cuda.current_stream())

    # ---------------------------------------------------------------- template-matching statistics form:
    if want("stats"):
        # the same batch updating per-pixel maximum / its template / sum / sum of squares in place of the maps
        bft.FwdImageInvFFTBatch(spec, intra_op=conj)  # the maps of every template, for the comparison
        torch.cuda.synchronize()
        pitch = 2 * (M // 2 + 1)
        maps = torch.stack([o[: pitch * M].view(M, pitch)[:, :M] for o in outs])
        want_mip, want_best = maps.max(dim=0)
        st = [torch.zeros(M * M, dtype=dt, device=dev) for dt in (torch.float32, torch.int32, torch.float32, torch.float32)]
        bft.SetBatchStats(*st, first_id=0)
        bft.SetStream(cap)
        g_st = capture(lambda: (bft.ResetBatchStats(), bft.FwdImageInvFFTBatch(spec, intra_op=conj)), 1)
        st_reps = max(2, K // 4)
        st_ms = time_graph(g_st, st_reps, warm=3) / st_reps
        bft.SetStream(torch.cuda.current_stream())
        torch.cuda.synchronize()
        scale = float(maps.abs().max())
        top2 = torch.topk(maps, 2, dim=0).values
        clear = (top2[0] - top2[1]) > 1e-5 * scale
        stats_2dtm = {
            "value": round(B * 1000.0 / st_ms, 1),
            "unit": "template correlations/s",
            "ms_per_call": round(st_ms, 5),
            "form": bft.Describe("batch").splitlines()[0],
            "mip_max_err_rel": float(f"{float((st[0].view(M, M) - want_mip).abs().max()) / scale:.3e}"),
            "best_agree_frac_clear_pixels": round(float((st[1].view(M, M) == want_best)[clear].float().mean()), 6),
            "sum_max_err_rel": float(f"{float((st[2].view(M, M) - maps.sum(0)).abs().max()) / float(maps.abs().sum(0).max()):.3e}"),
            "timing": f"HIP graph of ResetBatchStats + one FwdImageInvFFTBatch call in statistics form replayed {st_reps} times",
        }
        bft.SetBatchStats()
        del maps, want_mip, want_best, st, top2, clear

    # ---------------------------------------------------------------- split-BF16 form
    if want("split_bf16"):
        sft, souts, spk, s_ms, _ = batch_form("split_bf16")
        s_out0 = souts[0].cpu().numpy()
        s_peaks_same = bool(np.array_equal(spk.cpu().numpy(), b_pk_all))
        s_head = sft.Describe("batch").splitlines()[0]
        s_chunk = int(re.search(r"chunk=(\d+)", s_head).group(1))
        g_sx = capture(lambda: sft.RunBatchStage(2, spec, intra_op=conj), 1)
        s_xc_ms = time_graph(g_sx, reps) / reps / ((B + s_chunk - 1) // s_chunk)
        # matrix-core work of the split-BF16 correlation pass per launch: MFMA_PER_COLUMN
        # v_mfma_f32_32x32x16_bf16 (32 x 32 x 16 x 2 flops each) per template column, over the
        # fx / 2 column workgroups (the DC and Nyquist columns share one) and whole groups of 32 templates
        s_mfma_flops = round((s_chunk + 31) // 32 * 32 * (wf - 1) * MFMA_PER_COLUMN * 32 * 32 * 16 * 2)
        s_mfma_tflops = s_mfma_flops / (s_xc_ms * 1e-3) / 1e12
        sft.SetStream(torch.cuda.current_stream())

    # ---------------------------------------------------------------- uneven batch of 256 (96 x 128^2,
    if want("uneven"):
        # 96 x 256^2, 64 x 384^2), one call per step
        utm = W.batched_templates()
        usizes = [(t.shape[1], t.shape[0]) for t in utm]
        ubufs = [torch.from_numpy(W.to_padded(t).reshape(-1)).to(dev) for t in utm]
        uft = plan()
        uouts = [torch.zeros(uft.ReturnInvOutputMemorySize(), dtype=torch.float32, device=dev) for _ in utm]
        upk = torch.zeros(2 * len(utm), dtype=torch.int32, device=dev)
        uval = torch.zeros(len(utm), dtype=torch.float32, device=dev)
        uft.SetBatch(ubufs, usizes, uouts, upk, uval)
        uft.SetStream(cap)
        g_u = capture(lambda: uft.FwdImageInvFFTBatch(spec, intra_op=conj), 1)
        u_reps = max(2, K // 4)
        u_ms = time_graph(g_u, u_reps, warm=3) / u_reps
        uft.SetStream(torch.cuda.current_stream())
        uneven = {
            "templates": len(utm),
            "sizes": "96 x 128^2, 96 x 256^2, 64 x 384^2 (circular masks), 512^2 window",
            "value": round(len(utm) * 1000.0 / u_ms, 1),
            "ms_per_call": round(u_ms, 5),
            "form": uft.Describe("batch").splitlines()[0],
            "timing": f"HIP graph of one FwdImageInvFFTBatch call replayed {u_reps} times",
        }
        del uouts, ubufs
        # the headline's 384^2 templates, 256 per call (the 64 listed four times): the frame reduction
        # is shared by four times as many correlations
        lft = plan()
        louts = [torch.zeros(lft.ReturnInvOutputMemorySize(), dtype=torch.float32, device=dev) for _ in range(4 * B)]
        lpk = torch.zeros(8 * B, dtype=torch.int32, device=dev)
        lval = torch.zeros(4 * B, dtype=torch.float32, device=dev)
        lft.SetBatch([tbufs[i % B] for i in range(4 * B)], [(T, T)] * (4 * B), louts, lpk, lval)
        lft.SetStream(cap)
        g_l = capture(lambda: lft.FwdImageInvFFTBatch(spec, intra_op=conj), 1)
        l_ms = time_graph(g_l, u_reps, warm=3) / u_reps
        lft.SetStream(torch.cuda.current_stream())
        batch_256 = {
            "templates": 4 * B,
            "value": round(4 * B * 1000.0 / l_ms, 1),
            "ms_per_call": round(l_ms, 5),
            "form": lft.Describe("batch").splitlines()[0],
            "peaks_match_headline": bool(all(np.array_equal(c, b_pk_all.reshape(B, 2))
                                             for c in lpk.cpu().numpy().reshape(4, B, 2))),
            "timing": f"HIP graph of one FwdImageInvFFTBatch call replayed {u_reps} times",
        }
        del louts

    # ---------------------------------------------------------------- FP16 I/O: 256 FP16 templates of 512^2
    if want("config5"):
        # (radius-220 masks; 16 distinct, each listed 16 times) against the FP16 micrograph's __half2 spectrum,
        # FP16 512^2 outputs; FP32 and split-BF16 correlation passes
        T5, n5 = 512, 256
        mft16 = F.FourierTransformer("f16", "c32h", 2)
        mft16.SetForwardFFTPlan(N, N, 1, N, N, 1)
        mft16.SetInverseFFTPlan(N, N, 1, N, N, 1)
        spec16 = torch.from_numpy(W.to_padded(mic.astype(np.float16), dtype=np.float16).reshape(-1)).to(dev)
        mft16.FwdFFT(spec16, pre_op=F.Scale(1.0 / N))
        t5 = [torch.from_numpy(W.to_padded(W.masked_template(T5, 220.0, seed=500 + i).astype(np.float16),
                                           dtype=np.float16).reshape(-1)).to(dev) for i in range(16)]
        c5 = {}
        for prec in ("fp32", "split_bf16"):
            p5 = F.FourierTransformer("f16", "c32h", 2)
            p5.SetForwardFFTPlan(T5, T5, 1, N, N, 1)
            p5.SetInverseFFTPlan(N, N, 1, M, M, 1)
            p5.SetOutputWindow(-M // 2, -M // 2)
            o5 = [torch.zeros(p5.ReturnInvOutputMemorySize(), dtype=torch.float16, device=dev) for _ in range(n5)]
            pk5 = torch.zeros(2 * n5, dtype=torch.int32, device=dev)
            v5 = torch.zeros(n5, dtype=torch.float32, device=dev)
            p5.SetPrecision(prec)
            p5.SetBatch([t5[i % 16] for i in range(n5)], [(T5, T5)] * n5, o5, pk5, v5)
            p5.SetStream(cap)
            post5 = F.Scale(1.0 / (N * T5 * T5))
            g5 = capture(lambda: p5.FwdImageInvFFTBatch(spec16, intra_op=conj, post_op=post5), 1)
            r5 = max(2, K // 10)
            ms5 = time_graph(g5, r5, warm=3) / r5
            p5.SetStream(torch.cuda.current_stream())
            c5[prec] = {"value": round(n5 * 1000.0 / ms5, 1), "ms_per_call": round(ms5, 5),
                        "form": p5.Describe("batch").splitlines()[0], "peaks": pk5.cpu().numpy()}
            del o5
        # __half2 spectrum read once per call, each FP16 template in and each FP16 window out once
        bytes5_call = 33570816 + n5 * (524288 + 524288)
        config5 = {
            "templates": n5,
            "workload": "256 FP16 512^2 templates (radius-220 masks) vs the FP16 4096^2 micrograph's __half2 spectrum, "
                        "FP16 512^2 window outputs",
            "fp32": {k: v for k, v in c5["fp32"].items() if k != "peaks"},
            "split_bf16": {k: v for k, v in c5["split_bf16"].items() if k != "peaks"},
            "speedup_split_bf16_vs_fp32": round(c5["fp32"]["ms_per_call"] / c5["split_bf16"]["ms_per_call"], 3),
            "peaks_identical": bool(np.array_equal(c5["fp32"]["peaks"], c5["split_bf16"]["peaks"])),
            "compulsory_bytes_per_call": bytes5_call,
            "hbm_frac_per_call_fp32": round(bytes5_call / (c5["fp32"]["ms_per_call"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "timing": "HIP graph of one FwdImageInvFFTBatch call replayed; parity in tests/test_gpu_mfma.py",
        }
        del t5, spec16

    # ---------------------------------------------------------------- 3-D size change: padded forward
    if want("config4"):
        # 64^3 -> 256^3 and trimmed inverse 512^3 -> 128^3 (centred), each timed alone
        def time_3d(setup):
            p, fn = setup()
            p.SetStream(cap)
            with torch.cuda.stream(cap):
                fn()  # first call: workspace and twiddle table, outside the capture
            torch.cuda.synchronize()
            g3 = capture(fn, 10)  # ten transforms back to back per graph (steady state, as in `single`)
            r3 = max(2, K // 20)
            ms3 = time_graph(g3, r3) / (10 * r3)
            p.SetStream(torch.cuda.current_stream())
            return p, ms3

        def fwd3():
            p = F.FourierTransformer("f32", "c64", 3)
            p.SetForwardFFTPlan(64, 64, 64, 256, 256, 256)
            p.SetInverseFFTPlan(256, 256, 256, 256, 256, 256)
            vin = torch.from_numpy(W.to_padded(np.random.default_rng(6).standard_normal((64, 64, 64)).astype(np.float32))
                                   .reshape(-1)).to(dev)
            vout = torch.zeros(p.ReturnFwdOutputMemorySize(), dtype=torch.float32, device=dev)
            return p, lambda: p.FwdFFT(vin, vout)

        def inv3():
            p = F.FourierTransformer("f32", "c64", 3)
            p.SetForwardFFTPlan(512, 512, 512, 512, 512, 512)
            p.SetInverseFFTPlan(512, 512, 512, 128, 128, 128)
            p.SetOutputWindow(-64, -64, -64)
            sp3 = torch.randn(512 * 257 * 512 * 2, dtype=torch.float32, device=dev)
            o3 = torch.zeros(p.ReturnInvOutputMemorySize(), dtype=torch.float32, device=dev)
            return p, lambda: p.InvFFT(sp3, o3)

        _, f3_ms = time_3d(fwd3)
        _, i3_ms = time_3d(inv3)
        f3_bytes = 64 * 64 * 66 * 4 + 256 * 129 * 256 * 8
        i3_bytes = 512 * 257 * 512 * 8 + 128 * 128 * 130 * 4
        config4 = {
            "fwd_64_to_256": {"ms": round(f3_ms, 5), "compulsory_bytes": f3_bytes,
                              "hbm_frac": round(f3_bytes / (f3_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "inv_512_to_128_centred": {"ms": round(i3_ms, 5), "compulsory_bytes": i3_bytes,
                                       "hbm_frac": round(i3_bytes / (i3_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "timing": "HIP graph of ten transforms back to back, replayed; ms per transform; parity in "
                      "tests/test_gpu_3d.py",
        }

    # ---------------------------------------------------------------- single correlations
    if want("single"):
        ft = plan()
        stages = ft.Describe("xcorr").splitlines()[1:]
        out = torch.zeros(ft.ReturnInvOutputMemorySize(), dtype=torch.float32, device=dev)
        peak = torch.zeros(2, dtype=torch.int32, device=dev)
        pval = torch.zeros(1, dtype=torch.float32, device=dev)
        tbuf = torch.zeros(ft.ReturnInputMemorySize(), dtype=torch.float32, device=dev)
        tbuf.copy_(tbufs[0])

        def single_step():
            ft.FwdImageInvFFTArgMax(tbuf, spec, out, peak, pval, intra_op=conj)

        for _ in range(args.warmup):
            single_step()
        torch.cuda.synchronize()
        G = args.graph_batch
        ft.SetStream(cap)
        g_s = capture(single_step, G)
        s_replays = max(1, K // G)
        single_ms = time_graph(g_s, s_replays) / (s_replays * G)
        stage_ms = {}
        for i, name in enumerate(["r2c_x", "fused_y", "c2r_x"]):
            gp = capture(lambda i=i: ft.RunStages("xcorr", i, i, tbuf, spec, out, intra_op=conj), G)
            stage_ms[name] = time_graph(gp, s_replays) / (s_replays * G)
        ft.SetStream(torch.cuda.current_stream())
        fused_bytes = 8 * w * (T + N + M)

    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib * (1 << 20) // 4
        ca = torch.ones(nb, dtype=torch.float32, device=dev)
        cb = torch.empty_like(ca)
        cb.copy_(ca)
        torch.cuda.synchronize()
        del ca, cb

    # compulsory bytes of one step counted once per batch: the image spectrum read once, each template
    # read once and each 512^2 window written once
    step_bytes = 8 * w * N + B * (4 * T * T + 4 * M * M)
    value = B * 1000.0 / ms_step
    # FP32 arithmetic of the correlation pass per launch, counted as 5 n log2 n flops per n-point complex
    # transform (a forward and an inverse transform of fy points per template column) plus 6 flops per
    # point for the image product; one VALU lane does one FP32 FMA (2 flops) per cycle, so the
    # non-packed ceiling is half of the 157.3 TFLOP/s packed figure
    fft_flops = lambda n: 5.0 * n * np.log2(n)
    xc_flops = chunk * wf * (2 * fft_flops(fy) + 6.0 * fy)
    valu_tflops = xc_flops / (xc_ms * 1e-3) / 1e12
    result = {
        "metric": "XCorr/s 4096^2 vs 384^2 tmpl, trim 512^2",
        "value": round(value, 1),
        "unit": "correlations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"FwdImageInvFFTBatch: {B} masked 384x384 templates (zero padded) against one 4096x4096 "
                        "micrograph spectrum, each output trimmed to 512x512 lags -256..255, + argmax per template",
            "micrograph": N, "template": T, "output": M, "template_mask_radius": 160, "copies": 20,
            "batch": B, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_xcorr_col<512,2,0,float2> (batched correlation pass, {fx}x{fy} frame, {chunk} templates "
                      "per launch)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": TRAFFIC_PROFILE if traffic else None,
            "traffic_over_algorithmic": round(traffic / xc_bytes, 3) if traffic else None,
            "measured_hbm_frac": round(traffic / (xc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "algorithmic_bytes_per_launch": xc_bytes,
            "avg_launch_ms": round(xc_ms, 5),
            "launches_per_step": n_launch,
            "valu": {
                "flops_per_launch": round(xc_flops),
                "achieved_tflops": round(valu_tflops, 2),
                "peak_nonpacked_tflops": 78.6,
                "frac_nonpacked": round(valu_tflops / 78.6, 4),
                "peak_packed_tflops": 157.3,
                "counting": "5 n log2 n per n-point complex transform (forward + inverse per template column) "
                            "+ 6 per point for the image product",
            },
        },
        "step": {
            "form": head,
            "part_ms": {k: round(v, 5) for k, v in part_ms.items()},
            "set_batch_s": round(set_s, 3),
            "compulsory_bytes_per_step": step_bytes,
            "hbm_frac_per_step": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "compulsory_bytes_counting": "image spectrum read once per call + per template its 384^2 input "
                                         "and its 512^2 window output",
            "timing": f"HIP graph of one FwdImageInvFFTBatch call replayed {K} times after {args.warmup} warm-up "
                      "replays; part_ms from graphs of one part",
            "micrograph_fwd_ms": round(mic_fwd_ms, 4),
            "micrograph_fwd_hbm_frac": round(mic_fwd_bytes / (mic_fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "micrograph_fwd_timing": "ten FwdFFT calls on freshly copied 4096^2 micrographs enqueued back to back, "
                                     "HIP events around them, per transform",
            "peak_lag_template0": [b_pk[0] - M // 2, b_pk[1] - M // 2],
            "expected_peak_lag_template0": list(lags[0]),
        },
    }
    if want("split_bf16"):
        result["split_bf16"] = {
            "value": round(B * 1000.0 / s_ms, 1),
            "ms_per_step": round(s_ms, 5),
            "speedup_vs_fp32": round(ms_step / s_ms, 3),
            "peaks_identical_to_fp32": s_peaks_same,
            "form": s_head,
            "correlation_pass": {
                "kernel": MFMA_KERNEL,
                "avg_launch_ms": round(s_xc_ms, 5),
                "mfma_flops_per_launch": s_mfma_flops,
                "achieved_tflops": round(s_mfma_tflops, 1),
                "peak_bf16_dense_tflops": 2500.0,
                "mfma_frac": round(s_mfma_tflops / 2500.0, 4),
                "speedup_vs_fp32_pass": round(xc_ms / s_xc_ms, 3),
            },
        }
    if want("uneven"):
        result["uneven_batch"] = uneven
        result["batch_256"] = batch_256
    if want("stats"):
        result["stats_2dtm"] = stats_2dtm
    if want("config5"):
        result["config5_fp16"] = config5
    if want("config4"):
        result["config4_3d"] = config4
    if want("single"):
        result["single"] = {
            "value": round(1000.0 / single_ms, 1),
            "ms_per_corr": round(single_ms, 5),
            "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
            "fused_y_hbm_frac": round(fused_bytes / (stage_ms["fused_y"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "fused_y_algorithmic_bytes": fused_bytes,
            "timing": f"HIP graph of {G} fastfft_fwd_image_inv_argmax calls replayed {s_replays} times",
            "stages": stages,
        }

    if not args.no_cpu_baseline and want("cpu"):
        from oracle import cpu_baseline as CB
        from oracle import oracle as O

        # CPU baseline: the same correlation (template 0 against the micrograph, 512^2 window, argmax)
        # through the FFTW3 API on this host's cores, a bounded sample timed after two warm-up runs
        cb = CB.xcorr(mic, tmpl, M, warmup=2, min_corr=10, max_seconds=args.cpu_seconds)
        result["cpu_baseline"] = {
            "value": round(cb["corr_per_s"], 3),
            "unit": "correlations/s",
            "cores": cb["threads"],
            "kind": "port",
            "library": cb["library"],
            "cpu_model": cb["cpu_model"],
            "micrograph_fwd_ms": round(cb["mic_fwd_s"] * 1e3, 3),
            "peak_matches_gpu": [int(v) for v in cb["peak_xy"]] == b_pk,
            "sample": f"{cb['n_done']} FP32 correlations of template 0 (384^2 zero padded to 4096^2, r2c, "
                      f"image * conj, c2r, 512^2 crop, argmax) in {cb['corr_s']:.2f} s on {cb['threads']} threads, "
                      "after 2 untimed ones; the micrograph r2c is timed apart",
            "gpu_over_cpu": round(value / cb["corr_per_s"], 1) if cb["corr_per_s"] > 0 else None,
        }

        # parity: the FP64 oracle of the same correlation
        spec_h = spec.cpu().numpy()[: 2 * w * N].view(np.complex64).reshape(w, N).astype(np.complex128)
        ref64 = O.xcorr_2d(W.to_padded(tmpl), T, spec_h, N, N, M, M, -M // 2, -M // 2)
        wx, wy, _ = O.argmax_2d(ref64)

        def window(a):
            return a.reshape(M, F.padded_pitch(M))[:, :M].astype(np.float64)

        result["parity"] = {
            "rel_l2_vs_fp64": float(f"{O.rel_l2(window(b_out0), ref64):.3e}"),
            "max_abs_vs_fp64": float(f"{O.max_abs(window(b_out0), ref64):.3e}"),
            "max_abs_ref": float(f"{np.abs(ref64).max():.3e}"),
            "argmax_match": [wx, wy] == b_pk,
            "tolerance": 2e-6,
            "cpu_baseline_rel_l2_vs_fp64": float(f"{O.rel_l2(cb['window'].astype(np.float64), ref64):.3e}"),
        }
        if want("split_bf16"):
            result["split_bf16"]["rel_l2_vs_fp64"] = float(f"{O.rel_l2(window(s_out0), ref64):.3e}")
            result["split_bf16"]["tolerance"] = 2e-5
        if want("single"):
            result["single"]["rel_l2_vs_fp64"] = float(f"{O.rel_l2(window(out.cpu().numpy()), ref64):.3e}")
            result["single"]["argmax_match"] = [wx, wy] == [int(v) for v in peak.cpu().numpy()]
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
