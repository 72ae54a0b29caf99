"""Writes tests/golden/golden_small.npz: small input/expected-output vectors.

The reference stores no golden vectors (its tests are closed-form known-answer
tests), so these fixtures are generated here from first principles with
numpy only -- explicit DFT matrices in float64 for the tiny cases and
numpy.fft (pocketfft, float64) for the rest -- and never with the oracle they
are used to check.  Layout conventions are the reference's:
2-D spectra C[kx][ky] (include/FastFFT.cuh:750-767), 3-D spectra C[kz][kx][ky]
(src/tests/helper_functions.cuh:153-155), FFTW padded real rows
(include/FastFFT.h:363-404), unnormalised transforms, CONJ_MUL =
image * conj(template) (include/detail/functors.h:31-40).

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def pitch(x):
    return 2 * (x // 2 + 1)


def dft_mat(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def padded(img, x):
    out = np.zeros((img.shape[0], pitch(x)), np.float32)
    out[:, :x] = img
    return out


def brute_fwd_2d(img, n_x, n_y):
    """Zero pad to n_y x n_x, full 2-D DFT by matrices, keep kx <= n_x/2, transpose."""
    a = np.zeros((n_y, n_x))
    a[: img.shape[0], : img.shape[1]] = img
    F = dft_mat(n_y, -1) @ a @ dft_mat(n_x, -1)  # F[ky][kx]
    return F[:, : n_x // 2 + 1].T.copy()  # C[kx][ky]


def brute_inv_2d(spec, n_x, n_y, out_x, out_y, ox, oy):
    """Hermitian C2R along x of the y-inverse, by explicit sums; window at (ox, oy)."""
    cols = spec @ dft_mat(n_y, +1)  # [kx][y] after inverse along ky (symmetric matrix)
    h = n_x // 2
    out = np.zeros((out_y, out_x))
    for j in range(out_y):
        y = (oy + j) % n_y
        for i in range(out_x):
            x = (ox + i) % n_x
            v = cols[0, y].real + ((-1) ** x) * cols[h, y].real
            for k in range(1, h):
                v += 2.0 * (cols[k, y] * np.exp(2j * np.pi * k * x / n_x)).real
            out[j, i] = v
    return out


def numpy_inv_2d(spec, n_x, n_y, out_x, out_y, ox, oy):
    cols = np.fft.ifft(spec, axis=1) * n_y  # [kx][y]
    full = np.fft.irfft(cols.T, n=n_x, axis=1) * n_x  # [y][x]
    full = np.roll(full, (-oy, -ox), axis=(0, 1))
    return full[:out_y, :out_x]


def main():
    rng = np.random.default_rng(20261202)
    g = {}

    # 1. tiny brute force: 5x6 image (x=6, y=5) padded into 8x8, full DFT by matrices
    img = rng.standard_normal((5, 6)).astype(np.float32)
    g["bf_fwd_in"] = padded(img, 6)
    g["bf_fwd_dims"] = np.array([6, 5, 8, 8], np.int32)  # in_x, in_y, n_x, n_y
    g["bf_fwd_out"] = brute_fwd_2d(img.astype(np.float64), 8, 8)

    # 2. tiny brute force inverse with a wrapped window: random (non-Hermitian) C[kx][ky] 5x8
    spec = rng.standard_normal((5, 8)) + 1j * rng.standard_normal((5, 8))
    g["bf_inv_in"] = spec
    g["bf_inv_dims"] = np.array([8, 8, 4, 6, -2, -3], np.int32)  # n_x, n_y, out_x, out_y, ox, oy
    g["bf_inv_out"] = brute_inv_2d(spec, 8, 8, 4, 6, -2, -3)

    # 3. numpy: 24x20 (x=24, y=20) padded into 64x32
    img = rng.standard_normal((20, 24)).astype(np.float32)
    a = np.zeros((32, 64))
    a[:20, :24] = img
    g["np_fwd_in"] = padded(img, 24)
    g["np_fwd_dims"] = np.array([24, 20, 64, 32], np.int32)
    g["np_fwd_out"] = np.fft.rfft2(a).T.copy()

    # 4. numpy: full inverse then centred window 16x8 at (-8, -4) of a 64x32 spectrum
    spec = g["np_fwd_out"] * (1.0 + 0.25j)
    g["np_inv_in"] = spec
    g["np_inv_dims"] = np.array([64, 32, 16, 8, -8, -4], np.int32)
    g["np_inv_out"] = numpy_inv_2d(spec, 64, 32, 16, 8, -8, -4)

    # 5. cross-correlation: template 6x5 (x=6, y=5) against a 16x16 image, window 8x8 at (-4, -4)
    tmpl = rng.standard_normal((5, 6)).astype(np.float32)
    image = rng.standard_normal((16, 16)).astype(np.float32)
    ispec = np.fft.rfft2(image.astype(np.float64)).T.copy()
    t = np.zeros((16, 16))
    t[:5, :6] = tmpl
    tspec = np.fft.rfft2(t).T
    z = ispec * np.conj(tspec)
    g["xc_tmpl"] = padded(tmpl, 6)
    g["xc_image_spec"] = ispec
    g["xc_dims"] = np.array([6, 5, 16, 16, 8, 8, -4, -4], np.int32)  # tx, ty, n_x, n_y, out_x, out_y, ox, oy
    g["xc_out"] = numpy_inv_2d(z, 16, 16, 8, 8, -4, -4)
    # the same correlation directly in real space: c(s) = sum_x image(x + s) * t(x)
    direct = np.zeros((16, 16))
    for sy in range(16):
        for sx in range(16):
            direct[sy, sx] = np.sum(np.roll(image.astype(np.float64), (-sy, -sx), axis=(0, 1)) * t)
    g["xc_direct_full"] = direct

    # 6. 3-D: x=4, y=3, z=2 zero padded into 8x8x4
    vol = rng.standard_normal((2, 3, 4)).astype(np.float32)
    v = np.zeros((4, 8, 8))
    v[:2, :3, :4] = vol
    F = np.fft.rfftn(v)  # [kz][ky][kx]
    vp = np.zeros((2, 3, pitch(4)), np.float32)
    vp[..., :4] = vol
    g["f3_in"] = vp
    g["f3_dims"] = np.array([4, 3, 2, 8, 8, 4], np.int32)  # in_x, in_y, in_z, n_x, n_y, n_z
    g["f3_out"] = np.transpose(F, (0, 2, 1)).copy()  # C[kz][kx][ky]

    # 7. 3-D inverse with a wrapped window 4x2x2 at (-2, -1, -1) of an 8x8x4 spectrum
    spec3 = g["f3_out"] * (0.5 - 0.75j)
    cols = np.fft.ifft(np.fft.ifft(spec3, axis=0), axis=2) * 4 * 8  # [z][kx][y]
    full = np.fft.irfft(np.transpose(cols, (0, 2, 1)), n=8, axis=2) * 8  # [z][y][x]
    full = np.roll(full, (1, 1, 2), axis=(0, 1, 2))
    g["i3_in"] = spec3
    g["i3_dims"] = np.array([8, 8, 4, 4, 2, 2, -2, -1, -1], np.int32)
    g["i3_out"] = full[:2, :2, :4].copy()

    path = os.path.join(HERE, "golden_small.npz")
    np.savez_compressed(path, **g)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
