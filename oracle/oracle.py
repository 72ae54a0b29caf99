"""ctypes/numpy front end of the FP64 CPU oracle (oracle/fastfft_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the fastfft_amd product package.

Layouts follow the reference FourierTransformer (see fastfft_oracle.c header):
real arrays are FFTW padded ``[.., y][2*(x//2+1)]`` float32, 2-D spectra are
``C[kx][ky]``, 3-D spectra ``C[kz][kx][ky]``, transforms are unnormalised.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_c_int = ctypes.c_int
_c_double = ctypes.c_double


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_fft1d.argtypes = [_dp, _c_int, _c_int]
        L.oracle_fft1d.restype = None
        L.oracle_fwd_2d.argtypes = [_fp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _dp]
        L.oracle_fwd_2d.restype = _c_int
        L.oracle_inv_2d.argtypes = [_dp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _dp, _c_int]
        L.oracle_inv_2d.restype = _c_int
        L.oracle_xcorr_2d.argtypes = [_fp, _c_int, _c_int, _c_int, _dp, _c_int, _c_int, _c_double, _c_double,
                                      _c_int, _c_int, _c_int, _c_int, _c_double, _dp, _c_int]
        L.oracle_xcorr_2d.restype = _c_int
        L.oracle_fwd_3d.argtypes = [_fp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _dp]
        L.oracle_fwd_3d.restype = _c_int
        L.oracle_inv_3d.argtypes = [_dp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                    _c_double, _dp, _c_int]
        L.oracle_inv_3d.restype = _c_int
        L.oracle_argmax_2d.argtypes = [_dp, _c_int, _c_int, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                       ctypes.POINTER(_c_double)]
        L.oracle_argmax_2d.restype = None
        L.oracle_padded_memory_size.argtypes = [_c_int, _c_int, _c_int]
        L.oracle_padded_memory_size.restype = ctypes.c_long
        L.oracle_num_threads.argtypes = []
        L.oracle_num_threads.restype = _c_int
        _LIB = L
    return _LIB


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def padded_pitch(x: int) -> int:
    """Row pitch in floats of an FFTW padded real buffer (FastFFT.h:363-404)."""
    return 2 * (x // 2 + 1) if x % 2 == 0 else 2 * ((x - 1) // 2 + 1)


def padded_memory_size(x: int, y: int, z: int = 1) -> int:
    return int(lib().oracle_padded_memory_size(x, y, z))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def fft1d(z: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised complex FFT of a 1-D array (sign -1 forward, +1 inverse)."""
    a = np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).copy()
    v = a.view(np.float64)
    lib().oracle_fft1d(v, a.shape[0], sign)
    return a


def _as_padded_real(img: np.ndarray, x: int) -> np.ndarray:
    a = np.ascontiguousarray(img, dtype=np.float32)
    if a.shape[-1] < x:
        raise ValueError("row pitch smaller than logical x")
    return a


def fwd_2d(img: np.ndarray, in_x: int, n_x: int, n_y: int, pre_scale: float = 1.0) -> np.ndarray:
    """img: [in_y][pitch] float32.  Returns C[kx][ky] complex128, (n_x//2+1, n_y)."""
    a = _as_padded_real(img, in_x)
    in_y, pitch = a.shape
    out = np.zeros((n_x // 2 + 1, n_y), dtype=np.complex128)
    _check(lib().oracle_fwd_2d(a.reshape(-1), in_x, in_y, pitch, n_x, n_y, pre_scale, out.view(np.float64).reshape(-1)),
           "fwd_2d")
    return out


def inv_2d(spec: np.ndarray, n_x: int, n_y: int, out_x: int, out_y: int, ox: int = 0, oy: int = 0,
           post_scale: float = 1.0) -> np.ndarray:
    """spec: C[kx][ky] (n_x//2+1, n_y).  Returns [out_y][out_x] float64 window."""
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    assert s.shape == (n_x // 2 + 1, n_y), s.shape
    out = np.zeros((out_y, out_x), dtype=np.float64)
    _check(lib().oracle_inv_2d(s.view(np.float64).reshape(-1), n_x, n_y, out_x, out_y, ox, oy, post_scale,
                               out.reshape(-1), out_x), "inv_2d")
    return out


def xcorr_2d(tmpl: np.ndarray, tx: int, image_spec: np.ndarray, n_x: int, n_y: int, out_x: int, out_y: int,
             ox: int = 0, oy: int = 0, pre_scale: float = 1.0, intra_scale: float = 1.0,
             post_scale: float = 1.0) -> np.ndarray:
    """FwdImageInvFFT: window(Inv(image * conj(Fwd(pad(template)))))."""
    t = _as_padded_real(tmpl, tx)
    ty, tpitch = t.shape
    s = np.ascontiguousarray(image_spec, dtype=np.complex128)
    assert s.shape == (n_x // 2 + 1, n_y), s.shape
    out = np.zeros((out_y, out_x), dtype=np.float64)
    _check(lib().oracle_xcorr_2d(t.reshape(-1), tx, ty, tpitch, s.view(np.float64).reshape(-1), n_x, n_y, pre_scale,
                                 intra_scale, out_x, out_y, ox, oy, post_scale, out.reshape(-1), out_x), "xcorr_2d")
    return out


def fwd_3d(vol: np.ndarray, in_x: int, n_x: int, n_y: int, n_z: int, pre_scale: float = 1.0) -> np.ndarray:
    """vol: [in_z][in_y][pitch] float32.  Returns C[kz][kx][ky] complex128."""
    a = np.ascontiguousarray(vol, dtype=np.float32)
    in_z, in_y, pitch = a.shape
    out = np.zeros((n_z, n_x // 2 + 1, n_y), dtype=np.complex128)
    _check(lib().oracle_fwd_3d(a.reshape(-1), in_x, in_y, in_z, pitch, n_x, n_y, n_z, pre_scale,
                               out.view(np.float64).reshape(-1)), "fwd_3d")
    return out


def inv_3d(spec: np.ndarray, n_x: int, n_y: int, n_z: int, out_x: int, out_y: int, out_z: int, ox: int = 0,
           oy: int = 0, oz: int = 0, post_scale: float = 1.0) -> np.ndarray:
    """spec: C[kz][kx][ky].  Returns [out_z][out_y][out_x] float64 window."""
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    assert s.shape == (n_z, n_x // 2 + 1, n_y), s.shape
    out = np.zeros((out_z, out_y, out_x), dtype=np.float64)
    _check(lib().oracle_inv_3d(s.view(np.float64).reshape(-1), n_x, n_y, n_z, out_x, out_y, out_z, ox, oy, oz,
                               post_scale, out.reshape(-1), out_x), "inv_3d")
    return out


def argmax_2d(img: np.ndarray):
    """(x, y, value) of the maximum of a [ny][nx] float64 image, first hit wins."""
    a = np.ascontiguousarray(img, dtype=np.float64)
    ny, nx = a.shape
    ix, iy, v = _c_int(), _c_int(), _c_double()
    lib().oracle_argmax_2d(a.reshape(-1), nx, ny, nx, ctypes.byref(ix), ctypes.byref(iy), ctypes.byref(v))
    return ix.value, iy.value, v.value


def rel_l2(got: np.ndarray, want: np.ndarray) -> float:
    got = np.asarray(got, dtype=np.complex128 if np.iscomplexobj(got) else np.float64)
    want = np.asarray(want, dtype=got.dtype)
    den = np.linalg.norm(want.ravel())
    return float(np.linalg.norm((got - want).ravel()) / (den if den > 0 else 1.0))


def max_abs(got: np.ndarray, want: np.ndarray) -> float:
    return float(np.max(np.abs(np.asarray(got) - np.asarray(want)))) if np.size(want) else 0.0
