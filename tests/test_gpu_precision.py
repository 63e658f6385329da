"""GPU parity of the precision form: radix-16 stages of C2C launches on the matrix cores
(v_mfma_f32_16x16x16 f16 / bf16) with split operands and FP32 accumulation, against the FP64 oracle.

Modes: 1 = FP16 hi+lo with per-column power-of-two scaling, 2 = BF16 hi+lo, 3 = BF16 hi+mid+lo.
Tolerances (rel-L2 vs FP64): FP16 hi+lo and BF16 hi+mid+lo are held to the FP32 bar of 2e-6;
BF16 hi+lo keeps 16 significant bits per operand and is held to 5e-5 (its error is reported).
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, c2c, make, report, sign_of, to_dev

pytestmark = pytest.mark.gpu

MODES = {S.PREC_F16_SPLIT: ("f16x2", 2e-6), S.PREC_BF16_SPLIT: ("bf16x2", 5e-5),
         S.PREC_BF16_SPLIT3: ("bf16x3", 2e-6)}
MFMA_LENGTHS = [128, 256, 512, 1024, 2048, 4096]


def _stages(h):
    return h.describe()["stages"]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("n", MFMA_LENGTHS)
def test_mfma_leaf_lengths(torch_cuda, n, mode):
    name, tol = MODES[mode]
    batch = max(4, (1 << 18) // n) + 3  # a partial last tile
    x = oracle.uniform_complex((batch, n), 0x5EED0100 + n)
    for direction in (FWD, INV):
        y = c2c(x, [n], direction, precision=mode)
        rel, _ = report(f"{name} c2c n={n} b={batch} dir={direction:#x}", y, oracle.c2c(x, sign_of(direction)))
        assert rel <= tol


def test_mfma_stage_selected(torch_cuda):
    """The plan runs the MFMA kernels when a precision mode is set, and the plain ones otherwise;
    switching back to FP32 reproduces the plain result bit for bit."""
    import torch
    n, b = 4096, 64
    x = oracle.uniform_complex((b, n), 0x5EED0101)
    h = make([n], S.ASCEND_FFT_C2C, FWD, b)
    try:
        assert all(not s.get("mfma") for s in _stages(h))
        xd = to_dev(x)
        y0 = torch.empty_like(xd)
        assert h.exec_c2c(xd, y0) == 0
        assert h.set_precision(S.PREC_F16_SPLIT) == 0
        assert any(s.get("mfma") for s in _stages(h))
        y1 = torch.empty_like(xd)
        assert h.exec_c2c(xd, y1) == 0
        assert h.set_precision(S.PREC_FP32) == 0
        y2 = torch.empty_like(xd)
        assert h.exec_c2c(xd, y2) == 0
        assert h.synchronize() == 0
        assert torch.equal(y0, y2)
        ref = oracle.c2c(x)
        report("plain 4096x64", y0.cpu().numpy(), ref)
        rel, _ = report("f16x2 4096x64 (switched)", y1.cpu().numpy(), ref)
        assert rel <= 2e-6
    finally:
        h.destroy()


@pytest.mark.parametrize("mode", [S.PREC_F16_SPLIT, S.PREC_BF16_SPLIT3])
def test_mfma_headline_4096_b16384(torch_cuda, mode):
    name, tol = MODES[mode]
    n, b = 4096, 16384
    x = oracle.uniform_complex((b, n), 0x5EED0002)
    y = c2c(x, [n], FWD, precision=mode)
    rel, _ = report(f"{name} c2c 4096x16384", y, oracle.c2c(x))
    assert rel <= tol
    back = c2c(y, [n], INV, precision=mode) / n
    rel, _ = report(f"{name} roundtrip 4096x16384", back, x.astype(np.complex128))
    assert rel <= tol


@pytest.mark.parametrize("mode", list(MODES))
def test_c3_2d_1024x1024_b8_tones(torch_cuda, mode):
    name, tol = MODES[mode]
    x = oracle.tones_complex((8, 1024, 1024), seed=0x5EED0010)
    y = c2c(x, [1024, 1024], FWD, precision=mode)
    rel, _ = report(f"{name} 2d c2c 1024^2 b8 tones", y, oracle.c2c_nd(x, 2))
    assert rel <= tol


@pytest.mark.parametrize("mode", list(MODES))
def test_c3_3d_128cube_b4_tones(torch_cuda, mode):
    name, tol = MODES[mode]
    x = oracle.tones_complex((4, 128, 128, 128), seed=0x5EED0011)
    y = c2c(x, [128, 128, 128], FWD, precision=mode)
    rel, _ = report(f"{name} 3d c2c 128^3 b4 tones", y, oracle.c2c_nd(x, 3))
    assert rel <= tol


def test_f16_dynamic_range(torch_cuda):
    """Rows scaled from 1e-30 to 1e30: the per-column power-of-two scaling keeps every row at
    FP32-level relative error although FP16 alone spans only about 1e-8 .. 6.5e4."""
    n, b = 1024, 61
    x = oracle.uniform_complex((b, n), 0x5EED0102)
    scales = 10.0 ** np.linspace(-30, 30, b)
    xs = (x * scales[:, None]).astype(np.complex64)
    y = c2c(xs, [n], FWD, precision=S.PREC_F16_SPLIT)
    ref = oracle.c2c(xs)
    worst = 0.0
    for r in range(b):
        rel, _ = oracle.error(y[r], ref[r])
        worst = max(worst, rel)
    print(f"[parity] f16x2 dynamic range 1e-30..1e30: worst row rel-L2 {worst:.3e}", flush=True)
    assert worst <= 2e-6


@pytest.mark.parametrize("n", [65536, 2 ** 20, 1536 * 4])
def test_mfma_inside_four_step(torch_cuda, n):
    """Four-step plans whose leaves are MFMA lengths (column tiles and the fused twiddle)."""
    b = max(1, (1 << 21) // n)
    x = oracle.uniform_complex((b, n), 0x5EED0103 + n)
    y = c2c(x, [n], FWD, precision=S.PREC_F16_SPLIT)
    rel, _ = report(f"f16x2 four-step n={n}", y, oracle.c2c(x))
    assert rel <= 2e-6
