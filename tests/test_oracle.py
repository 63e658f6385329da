"""Pins the FP64 oracle before it is trusted (CPU only).

Checks, each chosen so that a plausible mistake (wrong sign, dropped term, transposed index,
wrong Hermitian fold) fails at least one of them:
  * the committed golden vectors (closed forms of the reference's example inputs);
  * closed forms: impulse, constant, pure tones, shifted impulses;
  * numpy.fft (pocketfft, an independent implementation) for 1D/2D/3D C2C, R2C, C2R;
  * the O(N^2) long-double DFT (a different algorithm) for many lengths;
  * Parseval and linearity;
  * PCG32 against the generator's published demo output (seed 42, stream 54).
"""
import numpy as np
import pytest

import oracle
from golden import load_fixture


def test_pcg32_known_answer():
    v = oracle.uniform(6, 42, 54).astype(np.float64)
    u = np.round((v + 1) / 2 * 2 ** 24).astype(np.int64)
    expected = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert list(u) == [e >> 8 for e in expected]


def test_uniform_range_and_determinism():
    a = oracle.uniform(100000, 7)
    assert a.min() >= -1.0 and a.max() < 1.0
    assert np.array_equal(a, oracle.uniform(100000, 7))
    assert abs(float(a.mean())) < 0.01
    g = oracle.gaussian(100000, 8)
    assert abs(float(g.std()) - 1.0) < 0.01


def test_golden_example_vectors():
    for name, fn in [("example_c2c_128x32", lambda x: oracle.c2c(x)),
                     ("example_r2c_256x32", lambda x: oracle.r2c(x)),
                     ("example_c2r_128x32_forward", lambda x: oracle.c2r(x, 128, oracle.FORWARD))]:
        fx = load_fixture(name)
        got = fn(fx["input"])
        ref = fx["output"]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (name, err)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 64, 97, 128, 1000, 1536, 2000, 4096, 997, 567])
def test_closed_forms(n):
    # impulse at 0 -> all ones; constant -> N delta; tone k0 -> N delta(k - k0); impulse at m -> w^{km}
    imp = np.zeros((1, n), np.complex128)
    imp[0, 0] = 1
    assert np.allclose(oracle.c2c(imp), np.ones((1, n)), atol=1e-14)
    const = np.ones((1, n), np.complex128)
    ref = np.zeros((1, n), np.complex128)
    ref[0, 0] = n
    assert np.allclose(oracle.c2c(const), ref, atol=1e-11 * n)
    k0 = (3 * n) // 7
    t = np.exp(2j * np.pi * k0 * np.arange(n) / n)[None, :]
    ref = np.zeros((1, n), np.complex128)
    ref[0, k0] = n
    assert np.allclose(oracle.c2c(t), ref, atol=1e-10 * n)
    # inverse direction flips the tone: e^{-2 pi i k0 n/N} -> N delta(k - k0) under the inverse
    assert np.allclose(oracle.c2c(np.conj(t), oracle.INVERSE), ref, atol=1e-10 * n)
    m = n // 3
    sh = np.zeros((1, n), np.complex128)
    sh[0, m] = 1
    assert np.allclose(oracle.c2c(sh), np.exp(-2j * np.pi * np.arange(n) * m / n)[None, :], atol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 6, 16, 45, 100, 243, 1024, 1536, 2000, 4096, 8192, 15000, 39601, 65536])
def test_vs_numpy(n):
    x = oracle.uniform_complex((3, n), n).astype(np.complex128)
    assert np.abs(oracle.c2c(x) - np.fft.fft(x)).max() < 1e-12 * max(1, np.sqrt(n))
    assert np.abs(oracle.c2c(x, oracle.INVERSE) - np.fft.ifft(x) * n).max() < 1e-12 * max(1, np.sqrt(n))
    xr = x.real.copy()
    assert np.abs(oracle.r2c(xr) - np.fft.rfft(xr)).max() < 1e-12 * max(1, np.sqrt(n))
    X = np.fft.rfft(xr)
    assert np.abs(oracle.c2r(X, n) - np.fft.irfft(X, n) * n).max() < 1e-11 * max(1, np.sqrt(n))


def test_c2r_drops_dc_and_nyquist_imag():
    n = 16
    X = np.zeros((1, n // 2 + 1), np.complex128)
    X[0, 0] = 1 + 5j
    X[0, n // 2] = 2 + 7j
    y = oracle.c2r(X, n)
    assert np.allclose(y, 1 + 2 * (-1.0) ** np.arange(n))


def test_multidim_vs_numpy():
    x = oracle.uniform_complex((2, 12, 10, 6), 11).astype(np.complex128)
    assert np.allclose(oracle.c2c_nd(x, 2), np.fft.fftn(x, axes=(-2, -1)), atol=1e-11)
    assert np.allclose(oracle.c2c_nd(x, 3), np.fft.fftn(x, axes=(-3, -2, -1)), atol=1e-11)
    xr = x.real.copy()
    assert np.allclose(oracle.r2c_nd(xr, 3), np.fft.rfftn(xr, axes=(-3, -2, -1)), atol=1e-11)
    X = np.fft.rfftn(xr, axes=(-3, -2, -1))
    y = oracle.c2r_nd(X, 6, 3)
    assert np.allclose(y, np.fft.irfftn(X, s=(12, 10, 6), axes=(-3, -2, -1)) * 720, atol=1e-9)


@pytest.mark.parametrize("n", list(range(1, 40)) + [64, 100, 127, 210, 1000, 1536])
def test_vs_long_double_dft(n):
    x = oracle.uniform_complex((1, n), 500 + n)[0]
    for sign in (oracle.FORWARD, oracle.INVERSE):
        ref = oracle.dft_ld(x, sign)
        got = oracle.c2c(x[None, :], sign)[0]
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_parseval_and_linearity():
    n = 2000
    x = oracle.uniform_complex((2, n), 1).astype(np.complex128)
    y = oracle.uniform_complex((2, n), 2).astype(np.complex128)
    X, Y = oracle.c2c(x), oracle.c2c(y)
    assert np.allclose(np.sum(np.abs(X) ** 2, -1), n * np.sum(np.abs(x) ** 2, -1), rtol=1e-12)
    assert np.allclose(oracle.c2c(2 * x - 3j * y), 2 * X - 3j * Y, atol=1e-10)


def test_error_report():
    ref = np.array([3.0 + 4.0j, 0.0], np.complex128)
    got = np.array([3.0 + 4.0j, 0.5], np.complex64)
    rel, mx = oracle.error(got, ref)
    assert abs(rel - 0.1) < 1e-12 and abs(mx - 0.5) < 1e-12


def test_fp32_baseline_accuracy():
    x = oracle.uniform_complex((64, 4096), 3)
    rel, _ = oracle.error(oracle.c2c_f32(x), oracle.c2c(x))
    assert rel < 1e-6


def test_planted_sign_mistake_is_caught():
    """A forward/inverse mix-up in the oracle would be caught by the numpy check above."""
    x = oracle.uniform_complex((1, 64), 5).astype(np.complex128)
    wrong = oracle.c2c(x, oracle.INVERSE)
    assert np.abs(wrong - np.fft.fft(x)).max() > 1e-3


# ---------------------------------------------------------------- ISTFT restatement

def _torch_cpu():
    torch = pytest.importorskip("torch")
    return torch


@pytest.mark.parametrize("n_fft,hop,frames,ch", [(16, 4, 9, 2), (64, 16, 12, 3), (400, 100, 7, 1), (101, 37, 5, 2)])
def test_istft_matches_torch_istft(n_fft, hop, frames, ch):
    """torch.istft is an independent implementation of the same definition (center=True,
    normalized=False, onesided=False, return_complex=True, length=None)."""
    torch = _torch_cpu()
    rng = np.random.default_rng(n_fft * 7 + hop)
    x = rng.standard_normal((ch, n_fft, frames)) + 1j * rng.standard_normal((ch, n_fft, frames))
    w = np.hanning(n_fft + 2)[1:-1] + 0.05  # strictly positive, so every sample has envelope
    ref = torch.istft(torch.from_numpy(x), n_fft=n_fft, hop_length=hop, win_length=n_fft,
                      window=torch.from_numpy(w), center=True, normalized=False, onesided=False,
                      return_complex=True).numpy()
    got = oracle.istft(x, n_fft, hop, w)
    assert got.shape == ref.shape == (ch, oracle.istft_out_len(n_fft, hop, frames))
    assert np.max(np.abs(got - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_istft_inverts_stft():
    """STFT (torch.stft, same framing) followed by the oracle ISTFT returns the signal."""
    torch = _torch_cpu()
    n_fft, hop = 256, 64
    rng = np.random.default_rng(5)
    sig = rng.standard_normal((2, 4000)) + 1j * rng.standard_normal((2, 4000))
    w = np.hanning(n_fft + 1)[:-1]
    spec = torch.stft(torch.from_numpy(sig), n_fft=n_fft, hop_length=hop, win_length=n_fft,
                      window=torch.from_numpy(w), center=True, onesided=False, return_complex=True).numpy()
    back = oracle.istft(spec, n_fft, hop, w)
    m = back.shape[1]
    assert m == oracle.istft_out_len(n_fft, hop, spec.shape[2])
    assert np.max(np.abs(back - sig[:, :m])) < 1e-10


def test_istft_closed_form_single_tone():
    """One frame pair with a rectangular window: the spectrum of a constant frame (n_fft at bin 0)
    gives back the constant wherever the frames overlap."""
    n_fft, hop, frames = 8, 4, 3
    x = np.zeros((1, n_fft, frames), dtype=np.complex128)
    x[0, 0, :] = n_fft * (2.0 - 1.0j)      # every frame is the constant 2 - i
    y = oracle.istft(x, n_fft, hop, np.ones(n_fft))
    assert y.shape == (1, oracle.istft_out_len(n_fft, hop, frames))
    assert np.allclose(y, 2.0 - 1.0j, atol=1e-13)
