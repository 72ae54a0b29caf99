"""3-D parity through the C ABI (Rank = 3 FourierTransformer).

Small shapes are compared element by element with the FP64 oracle.  The full
sizes (64^3 -> 256^3 padded forward, 512^3 -> 128^3 trimmed inverse)
are checked where the oracle would take minutes: against the transforms
evaluated directly in FP64 at sampled frequencies / output points (the DFT
sums are separable, so each sample costs one pass over the data), and the
trimmed inverse also against the untrimmed one cropped to the same window.
"""
import numpy as np
import pytest

from oracle import oracle as O

from . import gpu_helpers as H

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _vol(shape_zyx, seed):
    return np.random.default_rng(seed).standard_normal(shape_zyx).astype(np.float32)


def _fwd_gpu(v, n):
    """v: [z][y][x] float32 -> C[kz][kx][ky] complex128 from FwdFFT of the padded volume."""
    iz, iy, ix = v.shape
    ft = H.make_plan((ix, iy, iz), n, n, n, rank=3)
    a = H.to_dev(H.padded_real(v, ix).reshape(-1))
    out = H.dev_zeros(ft.ReturnFwdOutputMemorySize())
    ft.FwdFFT(a, out)
    return H.complex_from_f32(H.to_host(out), (n[2], n[0] // 2 + 1, n[1]))


@pytest.mark.parametrize("vin,n", [((16, 12, 10), (32, 32, 16)), ((8, 8, 8), (8, 8, 8)), ((20, 9, 30), (32, 16, 64)),
                                   ((64, 64, 64), (128, 128, 128)),
                                   # 256-point x-passes: real rows through the tile kernel, spectra stored transposed
                                   ((40, 20, 12), (256, 32, 16)), ((200, 16, 3), (256, 16, 8))])
def test_fwd_3d_vs_oracle(vin, n):
    H.log(f"3d fwd {vin} -> {n}")
    v = _vol(vin[::-1], seed=sum(vin))
    got = _fwd_gpu(v, n)
    want = O.fwd_3d(H.padded_real(v, vin[0]), vin[0], *n)
    assert O.rel_l2(got, want) < TOL


@pytest.mark.parametrize("n,nout,origin", [((32, 32, 16), (16, 8, 12), (-8, -4, -6)), ((16, 16, 16), (16, 16, 16), (0, 0, 0)),
                                            ((64, 32, 32), (20, 32, 8), (0, 0, 3)), ((128, 128, 128), (32, 32, 32), (-16, -16, -16))])
@pytest.mark.parametrize("order", ["default", "z_first"])
def test_inv_3d_vs_oracle(n, nout, origin, order, monkeypatch):
    """Trimmed 3-D inverses against the FP64 oracle: by default a trimmed y window puts the y-pass
    first (unit-stride lines of the whole input), FASTFFT_INV3D_Z_FIRST=1 the z-pass."""
    from fastfft_amd import workloads as W

    if order == "z_first":
        monkeypatch.setenv("FASTFFT_INV3D_Z_FIRST", "1")
    H.log(f"3d inv {n} -> {nout} at {origin} order={order}")
    rng = np.random.default_rng(n[0] + nout[0])
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((n[2], w, n[1])) + 1j * rng.standard_normal((n[2], w, n[1]))).astype(np.complex64)
    if n[0] == n[1] == n[2]:
        spec = W.hermitian_spectrum_3d(n[0], seed=3)
    ft = H.make_plan(n, n, n, nout, origin=origin, rank=3)
    a = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.InvFFT(a, out)
    got = H.real_window(H.to_host(out), *nout)
    want = O.inv_3d(spec, *n, *nout, *origin)
    assert O.rel_l2(got, want) < TOL


def test_round_trip_3d():
    n = (32, 32, 32)
    H.log(f"3d round trip {n}")
    v = _vol(n, seed=5)
    ft = H.make_plan(n, n, n, n, rank=3)
    buf = H.to_dev(H.padded_real(v, n[0]).reshape(-1))
    spec = H.dev_zeros(ft.ReturnFwdOutputMemorySize())
    ft.FwdFFT(buf, spec)
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.InvFFT(spec, out)
    got = H.real_window(H.to_host(out), *n)
    assert O.rel_l2(got, v.astype(np.float64) * np.prod(n)) < TOL


@pytest.mark.parametrize("tin,n,nout,origin", [((8, 8, 8), (32, 32, 32), (16, 16, 16), (-8, -8, -8)),
                                                ((12, 10, 6), (32, 16, 16), (32, 16, 16), (0, 0, 0))])
def test_xcorr_3d_vs_oracle(tin, n, nout, origin):
    H.log(f"3d xcorr {tin} in {n} -> {nout} at {origin}")
    t = _vol(tin[::-1], seed=11)
    img = _vol(n[::-1], seed=12)
    spec_img = O.fwd_3d(H.padded_real(img, n[0]), n[0], *n)
    ft = H.make_plan(tin, n, n, nout, origin=origin, rank=3)
    a = H.dev_zeros(max(ft.ReturnInputMemorySize(), ft.ReturnInvOutputMemorySize()))
    tp = H.padded_real(t, tin[0]).reshape(-1)
    a[: tp.size] = H.to_dev(tp)
    g = H.to_dev(H.complex_to_f32(spec_img))
    out = H.dev_poison(ft.ReturnInvOutputMemorySize())
    ft.FwdImageInvFFT(a, g, out)
    got = H.real_window(H.to_host(out), *nout)
    # the oracle's correlation = InvFFT(image * conj(FwdFFT(template)))
    spec_t = O.fwd_3d(H.padded_real(t, tin[0]), tin[0], *n)
    want = O.inv_3d(spec_img * np.conj(spec_t), *n, *nout, *origin)
    assert O.rel_l2(got, want) < TOL


# ------------------------------------------------------------------ full sizes
def test_full_fwd_64_to_256_sampled():
    H.log("full fwd 64^3 -> 256^3, sampled DFT")
    n = 256
    v = _vol((64, 64, 64), seed=4)
    got = _fwd_gpu(v, (n, n, n))
    rng = np.random.default_rng(44)
    S = 256
    kx = rng.integers(0, n // 2 + 1, S)
    ky = rng.integers(0, n, S)
    kz = rng.integers(0, n, S)
    x = np.arange(64)
    ex = np.exp(-2j * np.pi * np.outer(kx, x) / n)  # [S][x]
    ey = np.exp(-2j * np.pi * np.outer(ky, x) / n)
    ez = np.exp(-2j * np.pi * np.outer(kz, x) / n)
    t = v.astype(np.float64).reshape(64 * 64, 64) @ ex.T  # [z*y][S]
    t = t.reshape(64, 64, S)
    want = np.einsum("zys,sy,sz->s", t, ey, ez)
    assert O.rel_l2(got[kz, kx, ky], want) < TOL


def test_full_inv_512_to_128_sampled_and_cropped():
    from fastfft_amd import workloads as W

    H.log("full inv 512^3 -> 128^3 centred, sampled sums and untrimmed crop")
    n, m, o = 512, 128, -64
    spec = W.hermitian_spectrum_3d(n, seed=7)
    ft = H.make_plan((n, n, n), (n, n, n), (n, n, n), (m, m, m), origin=(o, o, o), rank=3)
    a = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.InvFFT(a, out)
    got = H.real_window(H.to_host(out), m, m, m)
    del out
    # untrimmed inverse on the GPU, cropped to the same wrapped window
    full = H.make_plan((n, n, n), (n, n, n), (n, n, n), (n, n, n), rank=3)
    fout = H.dev_zeros(full.ReturnInvOutputMemorySize())
    full.InvFFT(a, fout)
    fh = H.real_window(H.to_host(fout), n, n, n)
    del fout
    idx = (np.arange(m) + o) % n
    crop = fh[np.ix_(idx, idx, idx)]
    assert O.rel_l2(got, crop) < TOL
    # FP64 sums at sampled window points: y = Re sum_kz e_z sum_kx c_kx e_x sum_ky e_y C[kz][kx][ky]
    rng = np.random.default_rng(77)
    S = 12
    pts = rng.integers(0, m, (S, 3))  # (x, y, z) in window coordinates
    lag = (pts + o) % n
    k = np.arange(n)
    ey = np.exp(2j * np.pi * np.outer(k, lag[:, 1]) / n)  # [ky][S]
    kxs = np.arange(n // 2 + 1)
    ckx = np.where((kxs == 0) | (kxs == n // 2), 1.0, 2.0)
    ex = ckx[:, None] * np.exp(2j * np.pi * np.outer(kxs, lag[:, 0]) / n)  # [kx][S]
    ez = np.exp(2j * np.pi * np.outer(k, lag[:, 2]) / n)  # [kz][S]
    want = np.zeros(S, dtype=np.complex128)
    for z0 in range(0, n, 32):
        c = spec[z0:z0 + 32].astype(np.complex128)  # [kz][kx][ky]
        t = c.reshape(-1, n) @ ey  # [kz*kx][S]
        t = t.reshape(32, n // 2 + 1, S)
        u = np.einsum("zks,ks->zs", t, ex)
        want += np.einsum("zs,zs->s", u, ez[z0:z0 + 32])
    gv = got[pts[:, 2], pts[:, 1], pts[:, 0]]
    assert O.rel_l2(gv, want.real) < TOL


# ------------------------------------------------------------------ tiled z-passes (N = 256, 512)
@pytest.mark.parametrize("mode", ["default", "no_tile", "flags0", "flags3"])
@pytest.mark.parametrize("vin,n", [((20, 30, 40), (64, 64, 512)), ((32, 32, 32), (64, 64, 256))])
def test_fwd_3d_tiled_z_pass(vin, n, mode, monkeypatch):
    """Forward transforms whose z-pass (adjacent lines, far-apart elements) runs as tiles of 16 lines,
    against the FP64 oracle, with and without the tiled kernel, and with its tile order and 16-byte
    tile accesses forced off (flags0) and on (flags3)."""
    if mode == "no_tile":
        monkeypatch.setenv("FASTFFT_NO_TILE_LINES", "1")
    if mode.startswith("flags"):
        monkeypatch.setenv("FASTFFT_TILE_FLAGS", mode[-1])
    H.log(f"3d fwd {vin} -> {n}, z-pass mode={mode}")
    v = _vol(vin[::-1], seed=sum(vin) + 1)
    got = _fwd_gpu(v, n)
    want = O.fwd_3d(H.padded_real(v, vin[0]), vin[0], *n)
    assert O.rel_l2(got, want) < TOL


@pytest.mark.parametrize("mode", ["default", "no_tile", "z_first", "flags0", "flags3"])
def test_inv_3d_tiled_z_pass_window(mode, monkeypatch):
    """A trimmed inverse whose z-pass keeps 100 of 512 lags at an origin off the 64-lane grid (after the
    y-pass by default, before it with FASTFFT_INV3D_Z_FIRST=1; tile order and 16-byte tile accesses
    forced off and on with FASTFFT_TILE_FLAGS)."""
    if mode == "no_tile":
        monkeypatch.setenv("FASTFFT_NO_TILE_LINES", "1")
    if mode.startswith("flags"):
        monkeypatch.setenv("FASTFFT_TILE_FLAGS", mode[-1])
    if mode == "z_first":
        monkeypatch.setenv("FASTFFT_INV3D_Z_FIRST", "1")
    n, nout, origin = (128, 64, 512), (100, 60, 100), (-50, -30, -37)
    H.log(f"3d inv {n} -> {nout} at {origin}, z-pass mode={mode}")
    rng = np.random.default_rng(77)
    w = n[0] // 2 + 1
    spec = (rng.standard_normal((n[2], w, n[1])) + 1j * rng.standard_normal((n[2], w, n[1]))).astype(np.complex64)
    ft = H.make_plan(n, n, n, nout, origin=origin, rank=3)
    a = H.to_dev(H.complex_to_f32(spec))
    out = H.dev_zeros(ft.ReturnInvOutputMemorySize())
    ft.InvFFT(a, out)
    got = H.real_window(H.to_host(out), *nout)
    want = O.inv_3d(spec, *n, *nout, *origin)
    assert O.rel_l2(got, want) < TOL
