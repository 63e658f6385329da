"""GPU parity of the split-plane transforms (ASCEND_FFT_C2C_SEP + asdFftExecC2CSeparated):
real and imaginary parts in separate float32 arrays of one shape, 1D / 2D / 3D, against the FP64
oracle on the complex signal x_re + i x_im (tolerance 2e-6 rel-L2; 5e-6 for Bluestein lengths).

The reference documents signal lengths 2..256 for this entry point (docs/zh/FFT/FFT_1D.md:303-304)
and plans 1D and 3D only (fft_api.cpp:394-399, 451-458); this library accepts every length and
2D as well, through the same kernels as the interleaved transform.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, report, sign_of

pytestmark = pytest.mark.gpu


def _planes(x):
    import torch
    re = torch.from_numpy(np.ascontiguousarray(x.real.astype(np.float32))).cuda()
    im = torch.from_numpy(np.ascontiguousarray(x.imag.astype(np.float32))).cuda()
    return re, im


def _sep(x, dims, direction):
    import torch
    batch = x.size // int(np.prod(dims))
    h = S.FftHandle()
    try:
        if len(dims) == 1:
            st = h.make_plan_1d(dims[0], S.ASCEND_FFT_C2C_SEP, direction, batch)
        elif len(dims) == 2:
            st = h.make_plan_2d(dims[0], dims[1], S.ASCEND_FFT_C2C_SEP, direction, batch)
        else:
            st = h.make_plan_3d(dims[0], dims[1], dims[2], S.ASCEND_FFT_C2C_SEP, direction, batch)
        assert st == S.ACL_SUCCESS
        h.alloc_workspace()
        h.use_current_torch_stream()
        xr, xi = _planes(x)
        yr, yi = torch.empty_like(xr), torch.empty_like(xi)
        assert h.exec_c2c_separated(xr, xi, yr, yi) == S.ACL_SUCCESS
        assert h.synchronize() == S.ACL_SUCCESS
        assert np.array_equal(xr.cpu().numpy(), x.real.astype(np.float32)), "input was modified"
        return yr.cpu().numpy() + 1j * yi.cpu().numpy().astype(np.float64)
    finally:
        h.destroy()


@pytest.mark.parametrize("n", [2, 3, 16, 100, 128, 255, 256, 1, 1000, 4096, 997, 65536])
def test_separated_1d(torch_cuda, n):
    batch = max(1, min(32, (1 << 18) // n))
    x = oracle.uniform_complex((batch, n), 0x5EED0200 + n)
    tol = 5e-6 if n in (997, 255) else 2e-6
    for direction in (FWD, INV):
        y = _sep(x, [n], direction)
        rel, _ = report(f"c2c_sep n={n} b={batch} dir={direction:#x}", y, oracle.c2c(x, sign_of(direction)))
        assert rel <= tol


def test_separated_reference_example(torch_cuda):
    """The reference's split-plane example: x_re[i] = i, x_im[i] = i + 1 over batch 32 x N 256
    (docs/zh/FFT/FFT_1D.md:906-913)."""
    i = np.arange(32 * 256, dtype=np.float64)
    x = (i + 1j * (i + 1)).reshape(32, 256)
    y = _sep(x.astype(np.complex64), [256], FWD)
    rel, _ = report("c2c_sep example 256x32", y, oracle.c2c(x.astype(np.complex64)))
    assert rel <= 1e-6


@pytest.mark.parametrize("dims,batch", [([64, 64], 3), ([16, 32, 64], 2), ([128, 128, 128], 1), ([1024, 1024], 1)])
def test_separated_nd(torch_cuda, dims, batch):
    x = oracle.tones_complex(tuple([batch] + dims), seed=0x5EED0210 + len(dims))
    for direction in (FWD, INV):
        y = _sep(x, dims, direction)
        rel, _ = report(f"c2c_sep {dims} b={batch} dir={direction:#x}", y,
                        oracle.c2c_nd(x, len(dims), sign_of(direction)))
        assert rel <= 2e-6


def test_separated_validation(torch_cuda):
    import torch
    n, b = 256, 4
    h = S.FftHandle()
    try:
        assert h.make_plan_1d(n, S.ASCEND_FFT_C2C_SEP, FWD, b) == S.ACL_SUCCESS
        h.use_current_torch_stream()
        f = [torch.zeros((b, n), device="cuda") for _ in range(4)]
        c = torch.zeros((b, n), dtype=torch.complex64, device="cuda")
        # planes of different shapes: internal error, as in the reference (fft_api.cpp:1303-1330)
        short = torch.zeros((b, n - 1), device="cuda")
        assert h.exec_c2c_separated(f[0], f[1], f[2], short) == S.ACL_ERROR_INTERNAL_ERROR
        # complex planes, or a shape the plan does not describe: invalid parameter
        assert h.exec_c2c_separated(c, c, c, c) == S.ACL_ERROR_INVALID_PARAM
        g = [torch.zeros((b, 2 * n), device="cuda") for _ in range(4)]
        assert h.exec_c2c_separated(*g) == S.ACL_ERROR_INVALID_PARAM
        # the interleaved entry point does not run a split-plane plan
        assert h.exec_c2c(c, c) == S.ACL_ERROR_INVALID_PARAM
        assert h.exec_c2c_separated(*f) == S.ACL_SUCCESS
        assert h.synchronize() == S.ACL_SUCCESS
    finally:
        h.destroy()
