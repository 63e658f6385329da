"""FFT oracle (test infrastructure only).

The product library never imports this package. Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` use it, as the checker. It wraps ``oracle/fft_oracle.c``
(built into ``oracle/liboracle.so``): an FP64 Stockham FFT over the prime factorisation of N, an
O(N^2) long-double DFT, PCG32 seeded input generators, the rel-L2 / max-abs error report, and an
FP32 CPU FFT of the same staged algorithm (the CPU baseline).

Transform semantics restated from the reference (see the header of fft_oracle.c for file:line):
forward uses e^{-2 pi i kn/N}, inverse e^{+2 pi i kn/N} with no 1/N scaling; R2C returns bins
0..N/2; C2R transforms the Hermitian extension of N/2+1 bins and keeps the real part; 2D/3D
transforms are separable, R2C doing the real axis (last) first and C2R last.

``reference_steps`` / ``reference_istft_steps`` report which core and step list the reference
would build for an asdFft* call (``oracle/ref_rules.cpp``, built into ``oracle/libref_rules.so``).

Every function here is pinned by tests/test_oracle.py against closed forms (impulse, constant,
pure tones, the ramp the reference's examples feed in), numpy.fft, and the brute-force DFT; the
ISTFT restatement against torch.istft (an independent implementation of the same definition), STFT
round trips and a constant-frame closed form.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_RULES_PATH = os.path.join(_HERE, "libref_rules.so")
_lib = None
_rules = None

FORWARD = -1
INVERSE = +1


def build() -> str:
    """Compile liboracle.so from fft_oracle.c (gcc + OpenMP) and libref_rules.so from ref_rules.cpp."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _rules_lib():
    global _rules
    if _rules is None:
        if not os.path.exists(_RULES_PATH):
            build()
        L = ctypes.CDLL(_RULES_PATH)
        i64, sz = ctypes.c_int64, ctypes.c_size_t
        L.ref_steps_json.argtypes = [ctypes.c_int, ctypes.POINTER(i64), ctypes.c_int, ctypes.c_int, i64, ctypes.c_int,
                                     ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.ref_istft_steps_json.argtypes = [i64, i64, i64, i64, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        _rules = L
    return _rules


def _json_call(call):
    import json
    need = ctypes.c_size_t()
    call(None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    if call(buf, need.value, ctypes.byref(need)) != 0:
        raise RuntimeError("reference rules: buffer too small")
    return json.loads(buf.value.decode())


def reference_steps(dims, fft_type, batch, dim_type=0x10, direction=0x10):
    """The step list the reference builds for asdFftMakePlan{1,2,3}D with these arguments
    (core names and transposes, e.g. ["FftB(N=4096,batch=16384,radix=[64,64])"])."""
    dims = [int(d) for d in dims]
    arr = (ctypes.c_int64 * len(dims))(*dims)
    return _json_call(lambda buf, cap, need: _rules_lib().ref_steps_json(len(dims), arr, fft_type, direction, batch,
                                                                          dim_type, buf, cap, need))


def reference_istft_steps(channel, n_fft, frames, hop):
    """The reference's ISTFT step list (core/fft/istft.cpp:91-104)."""
    return _json_call(lambda buf, cap, need: _rules_lib().ref_istft_steps_json(channel, n_fft, frames, hop, buf, cap,
                                                                                need))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, dp, fp = ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
        L.oracle_fill_uniform.argtypes = [fp, i64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_fill_gaussian.argtypes = [fp, i64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_c2c_f64.argtypes = [dp, i64, i64, ctypes.c_int]
        L.oracle_c2c_axis_f64.argtypes = [dp, i64, i64, i64, ctypes.c_int]
        L.oracle_r2c_f64.argtypes = [dp, dp, i64, i64, ctypes.c_int]
        L.oracle_c2r_f64.argtypes = [dp, dp, i64, i64, ctypes.c_int]
        L.oracle_dft_ld.argtypes = [dp, dp, i64, ctypes.c_int]
        L.oracle_error_f32.argtypes = [fp, dp, i64, dp, dp]
        L.oracle_c2c_f32.argtypes = [fp, fp, i64, i64, ctypes.c_int, ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


# ------------------------------------------------------------------ seeded inputs

def uniform(count: int, seed: int, stream: int = 0) -> np.ndarray:
    """`count` float32 values uniform in [-1, 1) from PCG32(seed, stream)."""
    out = np.empty(int(count), dtype=np.float32)
    lib().oracle_fill_uniform(_fp(out), out.size, seed, stream)
    return out


def gaussian(count: int, seed: int, stream: int = 0) -> np.ndarray:
    """`count` float32 N(0,1) values (Box-Muller in FP64) from PCG32(seed, stream)."""
    out = np.empty(int(count), dtype=np.float32)
    lib().oracle_fill_gaussian(_fp(out), out.size, seed, stream)
    return out


def uniform_complex(shape, seed: int, stream: int = 0) -> np.ndarray:
    """complex64 array whose interleaved re/im floats are i.i.d. uniform[-1,1)."""
    shape = tuple(int(s) for s in np.atleast_1d(shape))
    n = int(np.prod(shape))
    return uniform(2 * n, seed, stream).view(np.complex64).reshape(shape)


def tones_complex(shape, seed: int, ntones: int = 3, amplitude: float = 1e3) -> np.ndarray:
    """Gaussian N(0,1) complex64 plus `ntones` pure tones of the given amplitude at seeded random
    bins of every transform (the precision-form input distribution). The tones are added over the
    full multi-dimensional index (one random frequency vector per tone, shared by the batch)."""
    shape = tuple(int(s) for s in shape)
    x = gaussian(2 * int(np.prod(shape)), seed).view(np.complex64).reshape(shape).astype(np.complex128)
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = shape[1:]
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    for _ in range(ntones):
        k = [int(rng.integers(0, d)) for d in dims]
        phase = sum(2.0 * np.pi * ki * g / d for ki, g, d in zip(k, grids, dims))
        x += amplitude * np.exp(1j * phase)[None, ...]
    return x.astype(np.complex64)


# ------------------------------------------------------------------ FP64 transforms

def c2c(x: np.ndarray, sign: int = FORWARD) -> np.ndarray:
    """FP64 C2C along the last axis (batch = all leading axes)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    _check(lib().oracle_c2c_f64(_dp(a), n, a.size // n if n else 0, sign), "c2c")
    return a


def c2c_axis(x: np.ndarray, axis: int, sign: int = FORWARD) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    axis = axis % a.ndim
    outer = int(np.prod(a.shape[:axis]))
    inner = int(np.prod(a.shape[axis + 1:]))
    _check(lib().oracle_c2c_axis_f64(_dp(a), outer, a.shape[axis], inner, sign), "c2c_axis")
    return a


def c2c_nd(x: np.ndarray, ndim: int, sign: int = FORWARD) -> np.ndarray:
    """C2C over the last `ndim` axes."""
    a = np.asarray(x, dtype=np.complex128)
    for ax in range(a.ndim - ndim, a.ndim):
        a = c2c_axis(a, ax, sign)
    return a


def istft_out_len(n_fft: int, hop: int, n_frames: int) -> int:
    """Output length of the reference's ISTFT with center=True and length 0: the overlap-add span
    n_fft + hop (n_frames - 1) minus n_fft/2 at each end (core/fft/istft.cpp:276-293)."""
    return n_fft + hop * (n_frames - 1) - 2 * (n_fft // 2)


def istft(x: np.ndarray, n_fft: int, hop: int, window: np.ndarray) -> np.ndarray:
    """ISTFT with the only parameters the reference runs (core/fft/istft.cpp:305-410 forces
    center=True, normalized=False, onesided=False, return_complex=True, length=0), in FP64.

    x: (channel, n_fft, n_frames) complex spectra. Per core/fft/fftcore/fft_core_istft_any.cpp:
    292-380: frames = IDFT_{n_fft}(x[c, :, t]) / n_fft; y[c, t hop + n] += frames[c, t, n] w[n];
    env[t hop + n] += w[n]^2; y / env, then the samples [n_fft/2, end - n_fft/2) are kept.
    Returns (channel, istft_out_len) complex128."""
    a = np.asarray(x, dtype=np.complex128)
    ch, nf, nfr = a.shape
    assert nf == n_fft
    w = np.asarray(window, dtype=np.float64)
    frames = c2c_axis(a, 1, INVERSE) / n_fft          # (ch, n_fft, n_frames)
    total = n_fft + hop * (nfr - 1)
    y = np.zeros((ch, total), dtype=np.complex128)
    env = np.zeros(total, dtype=np.float64)
    for t in range(nfr):
        y[:, t * hop:t * hop + n_fft] += frames[:, :, t] * w
        env[t * hop:t * hop + n_fft] += w * w
    start = n_fft // 2
    end = total - n_fft // 2
    return y[:, start:end] / env[start:end]


def r2c(x: np.ndarray, sign: int = FORWARD) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    n = a.shape[-1]
    batch = a.size // n
    out = np.empty(a.shape[:-1] + (n // 2 + 1,), dtype=np.complex128)
    _check(lib().oracle_r2c_f64(_dp(a), _dp(out), n, batch, sign), "r2c")
    return out


def c2r(x: np.ndarray, n: int, sign: int = INVERSE) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128)
    assert a.shape[-1] == n // 2 + 1
    batch = a.size // a.shape[-1]
    out = np.empty(a.shape[:-1] + (n,), dtype=np.float64)
    _check(lib().oracle_c2r_f64(_dp(a), _dp(out), n, batch, sign), "c2r")
    return out


def r2c_nd(x: np.ndarray, ndim: int, sign: int = FORWARD) -> np.ndarray:
    """R2C over the last `ndim` axes: real transform on the last axis, then C2C on the others."""
    a = r2c(x, sign)
    for ax in range(a.ndim - ndim, a.ndim - 1):
        a = c2c_axis(a, ax, sign)
    return a


def c2r_nd(x: np.ndarray, n_last: int, ndim: int, sign: int = INVERSE) -> np.ndarray:
    """C2R over the last `ndim` axes: C2C on the leading axes, then C2R on the last."""
    a = np.asarray(x, dtype=np.complex128)
    for ax in range(a.ndim - ndim, a.ndim - 1):
        a = c2c_axis(a, ax, sign)
    return c2r(a, n_last, sign)


def dft_ld(x: np.ndarray, sign: int = FORWARD) -> np.ndarray:
    """O(N^2) long-double DFT of a single 1D sequence (rounded to FP64 on return)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).ravel()
    out = np.empty_like(a)
    _check(lib().oracle_dft_ld(_dp(a), _dp(out), a.size, sign), "dft_ld")
    return out


# ------------------------------------------------------------------ error report

def error(got: np.ndarray, ref: np.ndarray):
    """(rel-L2, max-abs) of an FP32 result against an FP64 reference of the same shape."""
    g = np.ascontiguousarray(got)
    r = np.ascontiguousarray(ref)
    if np.iscomplexobj(g):
        g = g.astype(np.complex64, copy=False).view(np.float32)
    else:
        g = g.astype(np.float32, copy=False)
    if np.iscomplexobj(r):
        r = r.astype(np.complex128, copy=False).view(np.float64)
    else:
        r = r.astype(np.float64, copy=False)
    g = g.ravel()
    r = r.ravel()
    assert g.size == r.size, (g.size, r.size)
    rel = ctypes.c_double()
    mx = ctypes.c_double()
    lib().oracle_error_f32(_fp(g), _dp(r), g.size, ctypes.byref(rel), ctypes.byref(mx))
    return rel.value, mx.value


# ------------------------------------------------------------------ FP32 CPU baseline

def c2c_f32(x: np.ndarray, sign: int = FORWARD, threads: int = 0) -> np.ndarray:
    """FP32 C2C along the last axis with the same staged algorithm (CPU baseline)."""
    a = np.ascontiguousarray(x, dtype=np.complex64)
    n = a.shape[-1]
    out = np.empty_like(a)
    _check(lib().oracle_c2c_f32(_fp(a.view(np.float32)), _fp(out.view(np.float32)), n, a.size // n, sign, threads),
           "c2c_f32")
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
