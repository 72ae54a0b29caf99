#!/usr/bin/env python3
"""Regenerates the batch tuning cache (tuning.json beside libfastfft_amd.so) on the current GPU.

For each batch shape below, fastfft_set_batch times every candidate launch shape (templates per launch,
templates per group of the correlation pass, its sub-transform length) with FASTFFT_TUNE=1 and merges its
choice into the file named by FASTFFT_TUNING_WRITE (default: tuning.json beside the library; pass
--out to write elsewhere).  The shapes are those of the FastFFT correlation workloads:
  64 templates of 384^2 against a 4096^2 image, 512^2 centred window (FP32 and split-BF16);
  256 templates of 128^2 / 256^2 / 384^2 (96 / 96 / 64), same window;
  256 FP16 templates of 512^2 against a __half2 4096^2 spectrum, FP16 window (FP32 and split-BF16).
Timing runs use scratch outputs of the plan's, so the buffers below only need the right sizes.

Usage: python -m fastfft_amd.tune [--out PATH]
"""
from __future__ import annotations

import argparse
import os
import sys


def _shapes():
    n, m = 4096, 512
    yield "fp32 64 x 384^2", "f32", "c64", "fp32", [(384, 384)] * 64, n, m
    yield "split_bf16 64 x 384^2", "f32", "c64", "split_bf16", [(384, 384)] * 64, n, m
    yield "fp32 256 uneven", "f32", "c64", "fp32", [(128, 128)] * 96 + [(256, 256)] * 96 + [(384, 384)] * 64, n, m
    yield "fp16 256 x 512^2 fp32 pass", "f16", "c32h", "fp32", [(512, 512)] * 256, n, m
    yield "fp16 256 x 512^2 split_bf16 pass", "f16", "c32h", "split_bf16", [(512, 512)] * 256, n, m


def main() -> int:
    here = os.path.dirname(os.path.abspath(__file__))
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=os.path.join(here, "tuning.json"))
    args = ap.parse_args()
    os.environ["FASTFFT_TUNE"] = "1"
    os.environ["FASTFFT_TUNING_WRITE"] = os.path.abspath(args.out)
    sys.path.insert(0, os.path.dirname(here))
    import torch

    import fastfft_amd as F

    dev = torch.device("cuda:0")
    for name, it, ct, prec, sizes, n, m in _shapes():
        ft = F.FourierTransformer(it, ct, 2)
        lx, ly = max(s[0] for s in sizes), max(s[1] for s in sizes)
        ft.SetForwardFFTPlan(lx, ly, 1, n, n, 1)
        ft.SetInverseFFTPlan(n, n, 1, m, m, 1)
        ft.SetOutputWindow(-m // 2, -m // 2)
        dt = torch.float16 if it == "f16" else torch.float32
        distinct = {s: torch.zeros(F.padded_pitch(s[0]) * s[1], dtype=dt, device=dev) for s in set(sizes)}
        out = torch.zeros(ft.ReturnInvOutputMemorySize(), dtype=dt, device=dev)
        pk = torch.zeros(2 * len(sizes), dtype=torch.int32, device=dev)
        val = torch.zeros(len(sizes), dtype=torch.float32, device=dev)
        ft.SetPrecision(prec)
        ft.SetBatch([distinct[s] for s in sizes], sizes, [out] * len(sizes), pk, val)
        print(f"{name}: {ft.Describe('batch').splitlines()[0]}", flush=True)
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
