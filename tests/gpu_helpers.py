"""Helpers shared by the GPU parity tests (device buffers through torch, host oracle)."""
from __future__ import annotations

import sys

import numpy as np

import fastfft_amd as F


def log(msg: str) -> None:
    """Name the case before each GPU step, so a failing box still says which one ran."""
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


def torch_dev():
    import torch

    return torch, torch.device("cuda:0")


def dev_zeros(n_elems: int, dtype: str = "f32"):
    torch, d = torch_dev()
    td = {"f32": torch.float32, "f16": torch.float16}[dtype]
    return torch.zeros(int(n_elems), dtype=td, device=d)


def dev_poison(n_elems: int, dtype: str = "f32"):
    """An output buffer filled with NaN: whatever a transform leaves unwritten stays NaN, and a window
    that reads it fails the comparison."""
    torch, d = torch_dev()
    td = {"f32": torch.float32, "f16": torch.float16}[dtype]
    return torch.full((int(n_elems),), float("nan"), dtype=td, device=d)


def to_dev(a: np.ndarray):
    torch, d = torch_dev()
    return torch.from_numpy(np.ascontiguousarray(a)).to(d)


def to_host(t) -> np.ndarray:
    torch, _ = torch_dev()
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def padded_real(img: np.ndarray, x: int, dtype=np.float32) -> np.ndarray:
    """[.., y][x] -> FFTW padded [.., y][2*(x//2+1)]."""
    out = np.zeros(img.shape[:-1] + (F.padded_pitch(x),), dtype)
    out[..., :x] = img
    return out


def complex_from_f32(flat: np.ndarray, shape) -> np.ndarray:
    a = np.ascontiguousarray(flat[: 2 * int(np.prod(shape))], dtype=np.float32)
    return a.view(np.complex64).reshape(shape).astype(np.complex128)


def complex_to_f32(c: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(c.astype(np.complex64)).view(np.float32).reshape(-1)


def real_window(flat: np.ndarray, x: int, y: int, z: int = 1) -> np.ndarray:
    pitch = F.padded_pitch(x)
    a = np.asarray(flat[: pitch * y * z]).reshape(z, y, pitch)[..., :x].astype(np.float64)
    return a[0] if z == 1 else a


def make_plan(fwd_in, fwd_out, inv_in, inv_out, origin=(0, 0, 0), input_type="f32", image_type="c64", rank=2):
    ft = F.FourierTransformer(input_type, image_type, rank)
    ft.SetForwardFFTPlan(*fwd_in, *fwd_out)
    ft.SetInverseFFTPlan(*inv_in, *inv_out)
    if any(origin):
        ft.SetOutputWindow(*origin)
    return ft
