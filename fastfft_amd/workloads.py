"""Synthetic 2DTM-style workloads (micrographs with embedded template copies).

Seeds are fixed and documented so that every box, the GPU path and the CPU
oracle see identical input bits (numpy PCG64 is platform independent):

  round trip (64 x 64 into 512 x 512)   seed 1
  correlation micrograph / template      seed 2 (template), 3 (noise), 4 (positions)
  batched templates                      seed 5 + template index
  3-D volumes                            seed 6 (64^3 volume), 7 (512^3 spectrum)

A micrograph is N(0,1) noise plus copies of the template, scaled by an
amplitude, placed so that template pixel (x, y) lands on micrograph pixel
((x + dx) mod N, (y + dy) mod N).  With CONJ_MUL = image * conj(template) the
cross-correlation then peaks at lag +(dx, dy).  Copy 0 sits at a lag inside
the +-out/2 window centred on lag 0; the other copies are kept at least
out/2 + template size away from it, so the windowed argmax is a single clear
peak (SNR about 140 for a 384 x 384 template with a radius-160 mask).
"""
from __future__ import annotations

import numpy as np


def padded_pitch(x: int) -> int:
    return 2 * (x // 2 + 1) if x % 2 == 0 else 2 * ((x - 1) // 2 + 1)


def circular_mask(size: int, radius: float) -> np.ndarray:
    c = (size - 1) / 2.0
    yy, xx = np.mgrid[0:size, 0:size]
    return ((xx - c) ** 2 + (yy - c) ** 2) <= radius * radius


def masked_template(size: int, radius: float, seed: int, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    t = rng.standard_normal((size, size)).astype(np.float32)
    t *= circular_mask(size, radius)
    return t.astype(dtype)


def copy_lags(n_copies: int, n: int, window: int, tsize: int, seed: int) -> list[tuple[int, int]]:
    """Lag of every copy: copy 0 inside the centred window, the rest far outside it."""
    rng = np.random.default_rng(seed)
    half = window // 2
    lags = [(int(rng.integers(-half + 8, half - 8)), int(rng.integers(-half + 8, half - 8)))]
    keep_out = half + tsize
    while len(lags) < n_copies:
        dx, dy = (int(v) for v in rng.integers(-n // 2, n // 2, size=2))
        if max(abs(dx), abs(dy)) > keep_out:
            lags.append((dx, dy))
    return lags


def micrograph(n: int, template: np.ndarray, lags, amplitude: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((n, n)).astype(np.float32)
    ty, tx = template.shape
    t = template.astype(np.float32) * np.float32(amplitude)
    for dx, dy in lags:
        ys = (np.arange(ty) + dy) % n
        xs = (np.arange(tx) + dx) % n
        img[np.ix_(ys, xs)] += t
    return img


def to_padded(img: np.ndarray, dtype=np.float32) -> np.ndarray:
    """[.., y][x] -> FFTW padded [.., y][pitch]."""
    x = img.shape[-1]
    out = np.zeros(img.shape[:-1] + (padded_pitch(x),), dtype)
    out[..., :x] = img
    return out


def xcorr_config(n: int = 4096, tsize: int = 384, radius: float = 160.0, out: int = 512, n_copies: int = 20,
                 amplitude: float = 0.5):
    """The single-template correlation workload: (micrograph, template, copy lags)."""
    tmpl = masked_template(tsize, radius, seed=2)
    lags = copy_lags(n_copies, n, out, tsize, seed=4)
    mic = micrograph(n, tmpl, lags, amplitude, seed=3)
    return mic, tmpl, lags


def batched_templates(sizes_counts=((128, 96), (256, 96), (384, 64)), radius_frac: float = 160.0 / 384.0):
    """Templates of uneven sizes, each masked by a circle of radius radius_frac * size."""
    out = []
    idx = 0
    for size, count in sizes_counts:
        for _ in range(count):
            out.append(masked_template(size, radius_frac * size, seed=5 + idx))
            idx += 1
    return out


def hermitian_spectrum_3d(n: int, seed: int = 7) -> np.ndarray:
    """C[kz][kx][ky] (n, n//2+1, n) complex64 with N(0,1) re/im, made the exact
    half spectrum of a real volume: the kx = 0 and kx = n/2 planes satisfy
    X(kz, ky) = conj(X(-kz, -ky)) and their self-conjugate points are real."""
    rng = np.random.default_rng(seed)
    w = n // 2 + 1
    c = (rng.standard_normal((n, w, n)) + 1j * rng.standard_normal((n, w, n))).astype(np.complex64)
    for kx in (0, n // 2):
        plane = c[:, kx, :]  # [kz][ky]
        neg = np.conj(np.roll(np.flip(plane, axis=(0, 1)), 1, axis=(0, 1)))  # X(-kz, -ky)
        c[:, kx, :] = 0.5 * (plane + neg)
    return c
