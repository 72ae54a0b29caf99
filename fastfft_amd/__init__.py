"""fastfft_amd: MI355X-native FastFFT size-changing transforms (Python binding).

A ctypes binding over the C ABI in include/fastfft/fastfft.h, exposing the
reference's FourierTransformer interface (include/FastFFT.h:84-732) with the
same method names, argument meanings and layouts.  Device buffers are passed
as torch tensors (their data_ptr) or as raw integer device pointers, the idiom
the reference's binding experiments use (src/python/test_binding/test_pybind11.py:29,39).

The binding loads the in-tree libfastfft_amd.so and raises if it is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

__all__ = [
    "FastFFTError", "FourierTransformer", "Noop", "Scale", "ConjMul", "ConjMulThenScale", "lib", "library_path",
    "exported_symbols", "padded_pitch",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SUCCESS, INVALID_ARGUMENT, INVALID_DIMENSION, UNSUPPORTED, NOT_PLANNED, HIP_ERROR, OUT_OF_MEMORY, LDS_EXCEEDED = range(8)
_STATUS_NAMES = ["SUCCESS", "INVALID_ARGUMENT", "INVALID_DIMENSION", "UNSUPPORTED", "NOT_PLANNED", "HIP_ERROR",
                 "OUT_OF_MEMORY", "LDS_EXCEEDED"]

INPUT_F32, INPUT_F16 = 0, 1
IMAGE_C64, IMAGE_C32H = 0, 1
OP_NOOP, OP_SCALE, OP_CONJ_MUL, OP_CONJ_MUL_THEN_SCALE = 0, 1, 2, 3
PRECISION_FP32, PRECISION_SPLIT_BF16 = 0, 1


class FastFFTError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        name = _STATUS_NAMES[status] if 0 <= status < len(_STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {message}")


class _Op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("scale", ctypes.c_float)]


# The reference's functors (include/detail/functors.h:15-80): KernelFunction::my_functor
# <T, 0, NOOP>, <T, 2, SCALE>, <T, 4, CONJ_MUL>, <T, 4, CONJ_MUL_THEN_SCALE>.
def Noop() -> _Op:
    return _Op(OP_NOOP, 1.0)


def Scale(scale: float) -> _Op:
    return _Op(OP_SCALE, float(scale))


def ConjMul() -> _Op:
    return _Op(OP_CONJ_MUL, 1.0)


def ConjMulThenScale(scale: float) -> _Op:
    return _Op(OP_CONJ_MUL_THEN_SCALE, float(scale))


def library_path() -> str:
    return os.environ.get("FASTFFT_AMD_LIB", os.path.join(_HERE, "libfastfft_amd.so"))


def lib():
    """Load libfastfft_amd.so (built in-tree by fastfft_amd/build.py)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise FastFFTError(HIP_ERROR, f"{path} is not built; run `python fastfft_amd/build.py` "
                                      "(or __graft_entry__.build()) first")
    # A process that also uses torch must end up with torch's HIP runtime (torch ships its
    # own libamdhip64 / libhsa-runtime64): with this library's /opt/rocm runtime loaded
    # first, the library's HIP calls found no device once torch had initialised.  Loading
    # torch first makes both bind the same runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    P, I, V, S, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float
    sig = {
        "fastfft_create": [ctypes.POINTER(P), I, I, I],
        "fastfft_destroy": [P],
        "fastfft_set_fwd_plan": [P, S, S, S, S, S, S, I],
        "fastfft_set_inv_plan": [P, S, S, S, S, S, S, I],
        "fastfft_set_output_window": [P, I, I, I],
        "fastfft_set_stream": [P, V],
        "fastfft_set_precision": [P, ctypes.c_int],
        "fastfft_run_batch_stage": [P, ctypes.c_int, V, _Op, _Op, _Op],
        "fastfft_fwd": [P, V, V, _Op, _Op],
        "fastfft_inv": [P, V, V, _Op, _Op],
        "fastfft_fwd_image_inv": [P, V, V, V, _Op, _Op, _Op],
        "fastfft_fwd_image_inv_argmax": [P, V, V, V, _Op, _Op, _Op, V, V],
        "fastfft_fwd_image_inv_argmax_noalloc": [P, V, V, V, _Op, _Op, _Op, V, V],
        "fastfft_set_batch": [P, ctypes.c_int, V, V, V, V, V],
        "fastfft_fwd_image_inv_batch": [P, V, _Op, _Op, _Op],
        "fastfft_set_batch_stats": [P, V, V, V, V, I],
        "fastfft_reset_batch_stats": [P],
        "fastfft_wait": [P],
        "fastfft_copy_device_to_host_and_synchronize": [P, V, I],
        "fastfft_get_memory_sizes": [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)],
        "fastfft_get_dims": [P, I, ctypes.POINTER(I * 4)],
        "fastfft_get_buffer_pointer": [P, ctypes.POINTER(V)],
        "fastfft_workspace_bytes": [P, ctypes.POINTER(S)],
        "fastfft_argmax": [P, V, I, I, I, ctypes.POINTER(I * 2), ctypes.POINTER(F)],
        "fastfft_describe": [P, ctypes.c_char_p, ctypes.c_char_p, S],
        "fastfft_argmax_device": [P, V, I, I, I, V, V],
        "fastfft_run_stages": [P, ctypes.c_char_p, I, I, V, V, V, _Op, _Op, _Op],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.fastfft_last_error_string.argtypes = []
    L.fastfft_last_error_string.restype = ctypes.c_char_p
    L.fastfft_version.argtypes = []
    L.fastfft_version.restype = ctypes.c_char_p
    _LIB = L
    return L


def exported_symbols() -> list[str]:
    """Function names declared in include/fastfft/fastfft.h."""
    hdr = os.path.join(_HERE, "..", "include", "fastfft", "fastfft.h")
    with open(hdr) as f:
        text = f.read()
    decl = r"^\s*(?:fastfft_status|const\s+char\s*\*)\s*(fastfft_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.MULTILINE)))


def _check(status: int) -> None:
    if status != SUCCESS:
        msg = lib().fastfft_last_error_string().decode(errors="replace")
        raise FastFFTError(status, msg)


def padded_pitch(x: int) -> int:
    """Row pitch (elements) of an FFTW padded real buffer (FastFFT.h:363-404)."""
    return 2 * (x // 2 + 1) if x % 2 == 0 else 2 * ((x - 1) // 2 + 1)


def _ptr(buf) -> int | None:
    if buf is None:
        return None
    if isinstance(buf, int):
        return buf
    dp = getattr(buf, "data_ptr", None)
    if dp is None:
        raise TypeError("expected a torch tensor or an integer device pointer")
    return dp()


_TYPE_NAMES = {"f32": INPUT_F32, "float": INPUT_F32, "f16": INPUT_F16, "half": INPUT_F16}
_IMG_NAMES = {"c64": IMAGE_C64, "float2": IMAGE_C64, "c32h": IMAGE_C32H, "half2": IMAGE_C32H}


class FourierTransformer:
    """FourierTransformer<float, InputType, OtherImageType, Rank> over the C ABI.

    input_type: "f32" (float) or "f16" (__half); image_type: "c64" (float2) or
    "c32h" (__half2); rank: 2 or 3.  When torch is importable with a GPU, the
    plan runs on torch's current stream, so torch ops and transforms order
    naturally; pass stream=... to choose another (a raw hipStream_t int or a
    torch.cuda.Stream).
    """

    def __init__(self, input_type: str = "f32", image_type: str = "c64", rank: int = 2, stream=None):
        self._h = ctypes.c_void_p()
        _check(lib().fastfft_create(ctypes.byref(self._h), _TYPE_NAMES[input_type], _IMG_NAMES[image_type], int(rank)))
        self.input_type, self.image_type, self.rank = input_type, image_type, rank
        self._stream_set = False
        if stream is not None:
            self.SetStream(stream)

    # -- lifetime -------------------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().fastfft_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- plans ----------------------------------------------------------------------------
    def SetForwardFFTPlan(self, in_x, in_y, in_z, out_x, out_y, out_z, is_padded_output=True):
        _check(lib().fastfft_set_fwd_plan(self._h, in_x, in_y, in_z, out_x, out_y, out_z, int(bool(is_padded_output))))

    def SetInverseFFTPlan(self, in_x, in_y, in_z, out_x, out_y, out_z, is_padded_output=True):
        _check(lib().fastfft_set_inv_plan(self._h, in_x, in_y, in_z, out_x, out_y, out_z, int(bool(is_padded_output))))

    def SetOutputWindow(self, origin_x: int, origin_y: int, origin_z: int = 0):
        _check(lib().fastfft_set_output_window(self._h, int(origin_x), int(origin_y), int(origin_z)))

    def SetStream(self, stream) -> None:
        if stream is None:
            raw = None
        elif isinstance(stream, int):
            raw = stream
        else:
            raw = stream.cuda_stream
        _check(lib().fastfft_set_stream(self._h, raw))
        self._stream_set = True

    def SetPrecision(self, precision) -> None:
        """Arithmetic of the correlation pass: "fp32" (default) or "split_bf16" (matrix-core DFT-32
        stages with hi + lo BF16 operands where they apply; see fastfft_set_precision)."""
        code = {"fp32": PRECISION_FP32, "split_bf16": PRECISION_SPLIT_BF16}.get(precision, precision)
        _check(lib().fastfft_set_precision(self._h, int(code)))

    def _bind_torch_stream(self) -> None:
        if self._stream_set:
            return
        try:
            import torch

            if torch.cuda.is_available():
                self.SetStream(torch.cuda.current_stream())
                return
        except ImportError:
            pass
        self._stream_set = True

    # -- transforms -----------------------------------------------------------------------
    def FwdFFT(self, input_ptr, output_ptr=None, pre_op=None, intra_op=None):
        self._bind_torch_stream()
        _check(lib().fastfft_fwd(self._h, _ptr(input_ptr), _ptr(output_ptr), pre_op or Noop(), intra_op or Noop()))

    def InvFFT(self, input_ptr, output_ptr=None, intra_op=None, post_op=None):
        self._bind_torch_stream()
        _check(lib().fastfft_inv(self._h, _ptr(input_ptr), _ptr(output_ptr), intra_op or Noop(), post_op or Noop()))

    def FwdImageInvFFT(self, input_ptr, image_to_search, output_ptr=None, pre_op=None, intra_op=None, post_op=None):
        self._bind_torch_stream()
        _check(lib().fastfft_fwd_image_inv(self._h, _ptr(input_ptr), _ptr(image_to_search), _ptr(output_ptr),
                                           pre_op or Noop(), intra_op or ConjMul(), post_op or Noop()))

    def FwdImageInvFFTArgMax(self, input_ptr, image_to_search, output_ptr, peak_xy, value, pre_op=None, intra_op=None,
                             post_op=None):
        """Extension: FwdImageInvFFT plus the argmax of the output window, asynchronously.

        peak_xy: device int32[2] receiving (x, y) of the first maximum in row-major order;
        value: device float32[1] receiving its value.  2-D plans only.
        """
        self._bind_torch_stream()
        _check(lib().fastfft_fwd_image_inv_argmax(self._h, _ptr(input_ptr), _ptr(image_to_search), _ptr(output_ptr),
                                                  pre_op or Noop(), intra_op or ConjMul(), post_op or Noop(),
                                                  _ptr(peak_xy), _ptr(value)))

    def FwdImageInvFFTArgMaxNoAlloc(self, input_ptr, image_to_search, output_ptr, peak_xy, value, pre_op=None,
                                    intra_op=None, post_op=None):
        """Extension: FwdImageInvFFTArgMax whose intermediates live in the library's static device scratch
        (32 MiB per GPU): the call allocates no device memory.  Calls that overlap on one GPU race; plans
        that need more scratch raise FastFFTError (UNSUPPORTED).  2-D plans only."""
        self._bind_torch_stream()
        _check(lib().fastfft_fwd_image_inv_argmax_noalloc(self._h, _ptr(input_ptr), _ptr(image_to_search),
                                                          _ptr(output_ptr), pre_op or Noop(), intra_op or ConjMul(),
                                                          post_op or Noop(), _ptr(peak_xy), _ptr(value)))

    def SetBatch(self, templates, sizes, outputs, peaks_xy=None, values=None) -> None:
        """Extension: register templates (device buffers or pointers), their (x, y) sizes and output buffers
        for FwdImageInvFFTBatch; peaks_xy (device int32[2n]) and values (device float32[n]) are optional."""
        n = len(templates)
        if len(sizes) != n or len(outputs) != n:
            raise ValueError("templates, sizes and outputs must have the same length")
        arr_t = (ctypes.c_void_p * max(n, 1))(*[_ptr(t) for t in templates])
        arr_o = (ctypes.c_void_p * max(n, 1))(*[_ptr(o) for o in outputs])
        arr_s = (ctypes.c_int * max(2 * n, 1))(*[int(v) for s in sizes for v in s])
        self._batch_keepalive = (list(templates), list(outputs), peaks_xy, values)
        _check(lib().fastfft_set_batch(self._h, n, ctypes.cast(arr_t, ctypes.c_void_p), ctypes.cast(arr_s, ctypes.c_void_p),
                                       ctypes.cast(arr_o, ctypes.c_void_p), _ptr(peaks_xy), _ptr(values)))

    def FwdImageInvFFTBatch(self, image_to_search, pre_op=None, intra_op=None, post_op=None) -> None:
        """Extension: every correlation of the batch, asynchronously."""
        self._bind_torch_stream()
        _check(lib().fastfft_fwd_image_inv_batch(self._h, _ptr(image_to_search), pre_op or Noop(), intra_op or ConjMul(),
                                                 post_op or Noop()))

    def SetBatchStats(self, mip=None, best=None, sum=None, sum_sq=None, first_id: int = 0) -> None:
        """Extension (template matching, see fastfft_set_batch_stats): device maps of out_x * out_y
        elements (mip, sum, sum_sq float32, best int32) that FwdImageInvFFTBatch updates in place of the
        outputs: running per-pixel maximum, the template id (first_id + index) of its first occurrence,
        sum and sum of squares.  All None switches the statistics off."""
        self._stats_keepalive = (mip, best, sum, sum_sq)
        _check(lib().fastfft_set_batch_stats(self._h, _ptr(mip), _ptr(best), _ptr(sum), _ptr(sum_sq), int(first_id)))

    def ResetBatchStats(self) -> None:
        """Extension: mip = -inf, best = -1, sum = sum_sq = 0, asynchronously on the plan's stream."""
        self._bind_torch_stream()
        _check(lib().fastfft_reset_batch_stats(self._h))

    def RunBatchStage(self, stage: int, image_to_search, pre_op=None, intra_op=None, post_op=None) -> None:
        """One part of the batched call over the whole batch, for timing (see fastfft_run_batch_stage):
        0 frame reduction, 1 R2C rows, 2 correlation pass, 3 C2R rows, 4 argmax reduction."""
        self._bind_torch_stream()
        _check(lib().fastfft_run_batch_stage(self._h, int(stage), _ptr(image_to_search), pre_op or Noop(),
                                             intra_op or ConjMul(), post_op or Noop()))

    def Wait(self) -> None:
        _check(lib().fastfft_wait(self._h))

    def CopyDeviceToHostAndSynchronize(self, host_array, n_elements_to_copy: int = 0) -> None:
        """host_array: a writable numpy array (float32 or float16 by InputType)."""
        import numpy as np

        a = np.asarray(host_array)
        if not a.flags["C_CONTIGUOUS"] or not a.flags["WRITEABLE"]:
            raise ValueError("host array must be C-contiguous and writable")
        need = int(n_elements_to_copy) or None
        if need and a.size < need:
            raise ValueError("host array too small")
        _check(lib().fastfft_copy_device_to_host_and_synchronize(self._h, a.ctypes.data, int(n_elements_to_copy)))

    # -- sizes ----------------------------------------------------------------------------
    def _sizes(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().fastfft_get_memory_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def ReturnInputMemorySize(self) -> int:
        return self._sizes()[0]

    def ReturnFwdOutputMemorySize(self) -> int:
        return self._sizes()[1]

    def ReturnInvOutputMemorySize(self) -> int:
        return self._sizes()[2]

    def _dims(self, which: int):
        d = (ctypes.c_int * 4)()
        _check(lib().fastfft_get_dims(self._h, which, ctypes.byref(d)))
        return tuple(d)

    def ReturnFwdInputDimensions(self):
        return self._dims(0)

    def ReturnFwdOutputDimensions(self):
        return self._dims(1)

    def ReturnInvInputDimensions(self):
        return self._dims(2)

    def ReturnInvOutputDimensions(self):
        return self._dims(3)

    def GetDeviceBufferPointer(self) -> int:
        p = ctypes.c_void_p()
        _check(lib().fastfft_get_buffer_pointer(self._h, ctypes.byref(p)))
        return p.value or 0

    def WorkspaceBytes(self) -> int:
        n = ctypes.c_size_t()
        _check(lib().fastfft_workspace_bytes(self._h, ctypes.byref(n)))
        return n.value

    def Describe(self, op: str) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(lib().fastfft_describe(self._h, op.encode(), buf, len(buf)))
        return buf.value.decode()

    def RunStages(self, op: str, first: int, last: int, input_ptr, image_to_search=None, output_ptr=None, pre_op=None,
                  intra_op=None, post_op=None):
        """Run stages [first, last] of "fwd", "inv" or "xcorr" (see Describe)."""
        self._bind_torch_stream()
        intra = intra_op or (ConjMul() if op == "xcorr" else Noop())
        _check(lib().fastfft_run_stages(self._h, op.encode(), int(first), int(last), _ptr(input_ptr),
                                        _ptr(image_to_search), _ptr(output_ptr), pre_op or Noop(), intra,
                                        post_op or Noop()))

    def ArgMaxDevice(self, img, nx: int, ny: int, pitch: int, peak_xy, value) -> None:
        """Asynchronous argmax into device int32[2] peak_xy and float32[1] value."""
        self._bind_torch_stream()
        _check(lib().fastfft_argmax_device(self._h, _ptr(img), nx, ny, pitch, _ptr(peak_xy), _ptr(value)))

    def ArgMax(self, img, nx: int, ny: int, pitch: int):
        """(x, y, value) of the first maximum of a real [ny][pitch] device image."""
        self._bind_torch_stream()
        pk = (ctypes.c_int * 2)()
        v = ctypes.c_float()
        _check(lib().fastfft_argmax(self._h, _ptr(img), nx, ny, pitch, ctypes.byref(pk), ctypes.byref(v)))
        return pk[0], pk[1], v.value
