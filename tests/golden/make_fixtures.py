"""Writes the golden vectors in this directory (run: python tests/golden/make_fixtures.py).

The reference has no FFT golden data; its only deterministic fixtures are the example inputs
(example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:99-110: batch 32, N 128, x[i] = (i, i+1) over the
flattened tensor; example_acl_fft_r2c_1d.cpp:99-110: batch 32, N 256, x[i] = i;
example_acl_fft_c2r_1d.cpp:99-130: batch 32, N 128, bins (i, i+1), direction FORWARD).
The expected outputs here come from closed forms, not from the oracle:
  sum_n n w^n = N(N-1)/2 (k = 0) or N / (w - 1) (k != 0) with w = e^{-2 pi i k/N}, so for
  row b of the C2C input x[n] = (c+n)(1+i) + i, c = bN:  X[k] = (1+i)(cN delta_k + R_k) + iN delta_k;
and for the C2R example a direct long-double evaluation of the folded real sum
  y[n] = Re X0 + (-1)^n Re X_{N/2} + 2 sum_{k=1}^{N/2-1} (Re X_k cos t + Im X_k sin t), t = 2 pi k n/N
(forward direction: rotation rows of core/fft/fftcore/dft_c2r_core.cpp:90-105).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LD = np.longdouble


def ramp_dft(N):
    k = np.arange(N, dtype=LD)
    R = np.empty(N, dtype=np.clongdouble)
    R[0] = LD(N) * (N - 1) / 2
    ang = -2 * np.pi * k[1:] / N
    w = np.cos(ang) + 1j * np.sin(ang)
    R[1:] = N / (w - 1)
    return R


def c2c_example(N=128, B=32):
    x = np.empty((B, N), dtype=np.complex64)
    flat = np.arange(B * N, dtype=np.float64)
    x.real = flat.reshape(B, N)
    x.imag = flat.reshape(B, N) + 1
    R = ramp_dft(N)
    out = np.empty((B, N), dtype=np.clongdouble)
    for b in range(B):
        c = LD(b * N)
        X = (1 + 1j) * R
        X[0] += (1 + 1j) * c * N + 1j * N
        out[b] = X
    return x, out.astype(np.complex128)


def r2c_example(N=256, B=32):
    x = np.arange(B * N, dtype=np.float32).reshape(B, N)
    R = ramp_dft(N)[: N // 2 + 1]
    out = np.empty((B, N // 2 + 1), dtype=np.clongdouble)
    for b in range(B):
        X = R.copy()
        X[0] += LD(b * N) * N
        out[b] = X
    return x, out.astype(np.complex128)


def c2r_forward_example(N=128, B=32):
    H = N // 2 + 1
    x = np.empty((B, H), dtype=np.complex64)
    flat = np.arange(B * H, dtype=np.float64).reshape(B, H)
    x.real = flat
    x.imag = flat + 1
    n = np.arange(N, dtype=LD)
    out = np.empty((B, N), dtype=LD)
    for b in range(B):
        re = flat[b].astype(LD)
        im = re + 1
        y = re[0] + ((-1.0) ** n) * re[H - 1]
        for k in range(1, H - 1):
            t = 2 * np.pi * LD(k) * n / N
            y = y + 2 * (re[k] * np.cos(t) + im[k] * np.sin(t))
        out[b] = y
    return x, out.astype(np.float64)


def main():
    for name, (x, y) in {
        "example_c2c_128x32": c2c_example(),
        "example_r2c_256x32": r2c_example(),
        "example_c2r_128x32_forward": c2r_forward_example(),
    }.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), input=x, output=y)
        print(name, x.shape, y.shape)


if __name__ == "__main__":
    main()
