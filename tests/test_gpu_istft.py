"""GPU parity of the ISTFT (asdFftIstftMakePlan / asdFftExecIstft) against the FP64 oracle
restatement (oracle.istft, itself pinned against torch.istft and STFT round trips on the CPU).
Tolerance: rel-L2 2e-6 (5e-6 for Bluestein lengths); round trips 2e-6.

The reference runs one parameter set (center, not normalized, two-sided, complex output, no length
override; core/fft/istft.cpp:330-370) whatever the caller passes, and so does this library.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import report

pytestmark = pytest.mark.gpu


def _window(n, kind="hann+"):
    if kind == "hann+":
        return (np.hanning(n + 2)[1:-1] + 0.05).astype(np.float32)
    return (0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)).astype(np.float32)  # periodic Hann


def _istft(x, n_fft, hop, w, **kw):
    import torch
    ch, _, frames = x.shape
    h = S.FftHandle()
    try:
        xd = torch.from_numpy(np.ascontiguousarray(x.astype(np.complex64))).cuda()
        assert h.istft_make_plan(xd, n_fft, hop, n_fft, **kw) == S.ACL_SUCCESS
        h.alloc_workspace()
        h.use_current_torch_stream()
        wd = torch.from_numpy(np.ascontiguousarray(w)).cuda()
        y = torch.empty((ch, oracle.istft_out_len(n_fft, hop, frames)), dtype=torch.complex64, device="cuda")
        assert h.exec_istft(xd, wd, y) == S.ACL_SUCCESS
        assert h.synchronize() == S.ACL_SUCCESS
        return y.cpu().numpy()
    finally:
        h.destroy()


@pytest.mark.parametrize("ch,n_fft,hop,frames", [(1, 16, 4, 9), (2, 64, 16, 33), (3, 400, 100, 50),
                                                 (2, 512, 128, 200), (4, 1024, 256, 64), (2, 101, 37, 20),
                                                 (1, 1499, 300, 12), (2, 1000, 1000, 6)])
def test_istft_random_spectra(torch_cuda, ch, n_fft, hop, frames):
    rng = np.random.default_rng(n_fft * 131 + hop)
    x = (rng.uniform(-1, 1, (ch, n_fft, frames)) + 1j * rng.uniform(-1, 1, (ch, n_fft, frames))).astype(np.complex64)
    w = _window(n_fft)
    y = _istft(x, n_fft, hop, w)
    rel, _ = report(f"istft ch={ch} n_fft={n_fft} hop={hop} frames={frames}", y, oracle.istft(x, n_fft, hop, w))
    assert rel <= (5e-6 if n_fft == 1499 else 2e-6)


def test_istft_inverts_stft(torch_cuda):
    """A signal through torch.stft on the CPU (periodic Hann, hop n_fft/4) and back through the
    GPU ISTFT."""
    import torch
    n_fft, hop = 512, 128
    rng = np.random.default_rng(11)
    sig = rng.standard_normal((3, 20000)) + 1j * rng.standard_normal((3, 20000))
    w = _window(n_fft, "periodic")
    spec = torch.stft(torch.from_numpy(sig), n_fft=n_fft, hop_length=hop, win_length=n_fft,
                      window=torch.from_numpy(w.astype(np.float64)), center=True, onesided=False,
                      return_complex=True).numpy()
    y = _istft(spec.astype(np.complex64), n_fft, hop, w)
    rel, _ = report("istft(stft(x)) 3 x 20000", y, sig[:, :y.shape[1]])
    assert rel <= 2e-6


def test_istft_forced_parameters(torch_cuda):
    """center=False, normalized, onesided, length and real output are replaced by the one set the
    reference runs: the result equals the default call bit for bit."""
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((2, 64, 10)) + 1j * rng.standard_normal((2, 64, 10))).astype(np.complex64)
    w = _window(64)
    a = _istft(x, 64, 16, w)
    b = _istft(x, 64, 16, w, center=False, normalized=True, onesided=True, length=7, return_complex=False)
    assert np.array_equal(a, b)


def test_istft_exec_validation(torch_cuda):
    import torch
    ch, n_fft, hop, frames = 2, 64, 16, 10
    x = torch.zeros((ch, n_fft, frames), dtype=torch.complex64, device="cuda")
    w = torch.ones(n_fft, device="cuda")
    m = oracle.istft_out_len(n_fft, hop, frames)
    h = S.FftHandle()
    try:
        assert h.istft_make_plan(x, n_fft, hop, n_fft) == S.ACL_SUCCESS
        h.alloc_workspace()
        h.use_current_torch_stream()
        y = torch.empty((ch, m), dtype=torch.complex64, device="cuda")
        assert h.exec_istft(x, torch.ones(n_fft - 1, device="cuda"), y) == S.ACL_ERROR_INVALID_PARAM
        assert h.exec_istft(x, w, torch.empty((ch, m + 1), dtype=torch.complex64, device="cuda")) == \
            S.ACL_ERROR_INVALID_PARAM
        assert h.exec_istft(x, w, torch.empty((ch, m), device="cuda")) == S.ACL_ERROR_INVALID_PARAM
        x2 = torch.zeros((ch, n_fft, frames + 1), dtype=torch.complex64, device="cuda")
        assert h.exec_istft(x2, w, y) == S.ACL_ERROR_INVALID_PARAM
        # the interleaved entry point does not run an ISTFT plan
        assert h.exec_c2c(x, x) == S.ACL_ERROR_INVALID_PARAM
        assert h.exec_istft(x, w, y) == S.ACL_SUCCESS
        assert h.synchronize() == S.ACL_SUCCESS
        assert torch.count_nonzero(y).item() == 0
    finally:
        h.destroy()
