"""CPU baseline of the padded and trimmed cross-correlation (benchmark infrastructure).

Loaded only by the cpu_baseline leg of bench.py and by its CPU test; the product package never
imports it.  The pipeline (cpu_baseline.c) mirrors the reference's FFTW control
(src/fastfft/Image.cu:116-170): micrograph r2c once, then per template zero-pad, r2c, image * conj,
c2r, crop the wrapped window, argmax.  The library is resolved at run time, in this order:
  1. FFTW (libfftw3f.so.3 + libfftw3f_threads.so.3),
  2. Intel MKL's FFTW3 interface (libmkl_rt, searched in MKL_DIRS),
  3. scipy.fft (pocketfft) with workers = the thread count,
and reported by name and version with the thread count and the CPU model.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MKL_DIRS = ["/opt/conda/lib", "/opt/intel/oneapi/mkl/latest/lib/intel64", "/usr/lib/x86_64-linux-gnu"]
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "libcpu_baseline.so"], check=True)
    return os.path.join(_HERE, "libcpu_baseline.so")


def threads() -> int:
    """Host threads this process may use: the CPU affinity, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcpu_baseline.so")
        if not os.path.exists(path):
            build()
        # MKL's threading layer on the GNU OpenMP runtime the library itself uses
        os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
        L = ctypes.CDLL(path)
        L.cpub_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.cpub_open.restype = ctypes.c_int
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.cpub_xcorr.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, dp, dp, ip, fp, ip]
        L.cpub_xcorr.restype = ctypes.c_int
        _LIB = L
    return _LIB


def library() -> tuple[int, str]:
    """(kind, description): 1 FFTW, 2 MKL FFTW3 API, 0 neither (scipy.fft is used)."""
    L = _lib()
    desc = ctypes.create_string_buffer(256)
    for d in MKL_DIRS:
        k = L.cpub_open(d.encode(), desc, len(desc))
        if k:
            return k, desc.value.decode()
    return 0, "none"


def _xcorr_scipy(mic, tmpl, m, nthreads, warmup, min_corr, max_seconds):
    import scipy
    import scipy.fft as sf

    n, t = mic.shape[0], tmpl.shape[0]
    t0 = time.perf_counter()
    img = sf.rfft2(mic, workers=nthreads)
    mic_fwd = time.perf_counter() - t0
    pad = np.zeros((n, n), np.float32)
    done, tc, it = 0, 0.0, 0
    lag = (np.arange(m) - m // 2) % n
    while True:
        timed = it >= warmup
        if timed and done >= min_corr and tc >= max_seconds:
            break
        ts = time.perf_counter()
        pad[:t, :t] = tmpl
        spec = sf.rfft2(pad, workers=nthreads)
        spec = img * np.conj(spec)
        full = sf.irfft2(spec, s=(n, n), workers=nthreads) * np.float32(n * n)  # unnormalised like FFTW
        window = full[np.ix_(lag, lag)]
        k = int(np.argmax(window))
        pad[:t, :t] = 0
        if timed:
            tc += time.perf_counter() - ts
            done += 1
        it += 1
    return {"mic_fwd_s": mic_fwd, "corr_s": tc, "n_done": done, "window": window.astype(np.float32),
            "peak_xy": (k % m, k // m), "library": f"scipy.fft {scipy.__version__} (pocketfft)"}


def xcorr(mic: np.ndarray, tmpl: np.ndarray, m: int, nthreads: int | None = None, warmup: int = 2,
          min_corr: int = 10, max_seconds: float = 10.0) -> dict:
    """Times the CPU pipeline on the n x n micrograph and t x t template (float32, row-major): warmup
    untimed correlations, then at least min_corr and at least max_seconds of timed ones."""
    mic = np.ascontiguousarray(mic, dtype=np.float32)
    tmpl = np.ascontiguousarray(tmpl, dtype=np.float32)
    nthreads = nthreads or threads()
    kind, desc = library()
    if kind == 0:
        r = _xcorr_scipy(mic, tmpl, m, nthreads, warmup, min_corr, max_seconds)
    else:
        L = _lib()
        n, t = mic.shape[0], tmpl.shape[0]
        window = np.zeros((m, m), np.float32)
        mf, cs = ctypes.c_double(), ctypes.c_double()
        nd = ctypes.c_int()
        pk = (ctypes.c_int * 2)()
        rc = L.cpub_xcorr(mic.reshape(-1), tmpl.reshape(-1), n, t, m, nthreads, warmup, min_corr, max_seconds,
                          ctypes.byref(mf), ctypes.byref(cs), ctypes.byref(nd), window.reshape(-1), pk)
        if rc != 0:
            raise RuntimeError(f"cpu baseline failed with code {rc}")
        r = {"mic_fwd_s": mf.value, "corr_s": cs.value, "n_done": nd.value, "window": window,
             "peak_xy": (pk[0], pk[1]), "library": desc}
    r["threads"] = nthreads
    r["cpu_model"] = cpu_model()
    r["corr_per_s"] = r["n_done"] / r["corr_s"] if r["corr_s"] > 0 else 0.0
    return r
