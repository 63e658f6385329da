"""GPU parity of 2D / 3D and vertical (strided) transforms against the FP64 oracle.

The precision-form workloads (2D C2C 1024x1024 batch 8, 3D C2C 128^3 batch 4; Gaussian N(0,1)
plus 3 tones of amplitude 1e3) are run here through the plain FP32 form; rel-L2 <= 2e-6 and the
max-abs error are reported. Round trips check inverse(forward(x)) = N x.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, c2c, c2r, r2c, report, sign_of

pytestmark = pytest.mark.gpu
TOL = 2e-6


def test_c3_2d_1024x1024_b8_tones(torch_cuda):
    x = oracle.tones_complex((8, 1024, 1024), seed=0x5EED0010)
    y = c2c(x, [1024, 1024], FWD)
    rel, mx = report("2d c2c 1024^2 b8 tones", y, oracle.c2c_nd(x, 2))
    assert rel <= TOL
    back = c2c(y, [1024, 1024], INV) / (1024 * 1024)
    rel, _ = report("2d roundtrip", back, x.astype(np.complex128))
    assert rel <= TOL


def test_c3_3d_128cube_b4_tones(torch_cuda):
    x = oracle.tones_complex((4, 128, 128, 128), seed=0x5EED0011)
    y = c2c(x, [128, 128, 128], FWD)
    rel, mx = report("3d c2c 128^3 b4 tones", y, oracle.c2c_nd(x, 3))
    assert rel <= TOL
    back = c2c(y, [128, 128, 128], INV) / 128 ** 3
    rel, _ = report("3d roundtrip", back, x.astype(np.complex128))
    assert rel <= TOL


@pytest.mark.parametrize("dims,batch", [((64, 64), 2), ((100, 200), 3), ((2048, 3000), 1), ((7, 13), 5),
                                        ((4096, 33), 1), ((1, 128), 4), ((128, 1), 4), ((17, 31), 2),
                                        ((32, 48, 20), 2), ((8, 8, 8), 7), ((5, 1024, 3), 1), ((1, 1, 64), 2),
                                        ((30, 997, 2), 1)])
def test_c2c_nd(torch_cuda, dims, batch):
    x = oracle.uniform_complex((batch,) + dims, seed=sum(dims) + batch)
    tol = 5e-6 if any(d in (17, 31, 997) for d in dims) else TOL
    for direction in (FWD, INV):
        y = c2c(x, list(dims), direction)
        rel, _ = report(f"c2c nd {dims} b{batch} dir={direction:#x}", y, oracle.c2c_nd(x, len(dims), sign_of(direction)))
        assert rel <= tol


@pytest.mark.parametrize("dims,batch", [((64, 64), 2), ((100, 200), 3), ((128, 128), 2), ((1024, 4096), 2),
                                        ((33, 17), 2), ((32, 48, 20), 2), ((16, 16, 15), 3), ((7, 9, 2), 2),
                                        ((64, 2, 65536), 1)])
def test_r2c_c2r_nd(torch_cuda, dims, batch):
    nd = len(dims)
    x = oracle.uniform(batch * int(np.prod(dims)), seed=7 + sum(dims)).reshape((batch,) + dims)
    tol = 5e-6 if 17 in dims or 33 in dims else TOL
    X = r2c(x, list(dims))
    rel, _ = report(f"r2c nd {dims}", X, oracle.r2c_nd(x, nd))
    assert rel <= tol
    Xh = oracle.r2c_nd(x, nd).astype(np.complex64)
    y = c2r(Xh, list(dims))
    rel, _ = report(f"c2r nd {dims}", y, oracle.c2r_nd(Xh, dims[-1], nd))
    assert rel <= tol
    back = y / np.prod(dims)
    rel, _ = report(f"r2c/c2r roundtrip nd {dims}", back, x.astype(np.float64))
    assert rel <= tol


def test_reference_2d_examples(torch_cuda):
    """Inputs of example/A2/FFT/2d: C2C 64x64 batch 2 with x[i] = (i, i+1); R2C 1024x4096 batch 2 with
    x[i] = i; C2R 128x128 batch 2 with bins (i, i+1), direction FORWARD."""
    n = 2 * 64 * 64
    x = (np.arange(n) + 1j * (np.arange(n) + 1)).astype(np.complex64).reshape(2, 64, 64)
    y = c2c(x, [64, 64], FWD)
    rel, _ = report("example 2d c2c 64x64", y, oracle.c2c_nd(x, 2))
    assert rel <= 1e-6
    xr = np.arange(2 * 1024 * 4096, dtype=np.float32).reshape(2, 1024, 4096)
    X = r2c(xr, [1024, 4096])
    rel, _ = report("example 2d r2c 1024x4096", X, oracle.r2c_nd(xr, 2))
    assert rel <= TOL
    h = 65
    xc = (np.arange(2 * 128 * h) + 1j * (np.arange(2 * 128 * h) + 1)).astype(np.complex64).reshape(2, 128, h)
    yr = c2r(xc, [128, 128], FWD)
    rel, _ = report("example 2d c2r 128x128 forward", yr, oracle.c2r_nd(xc, 128, 2, oracle.FORWARD))
    assert rel <= TOL


@pytest.mark.parametrize("n,batch", [(4096, 128), (256, 1000), (65536, 128), (1536, 40), (997, 8), (8, 3)])
def test_vertical_1d(torch_cuda, n, batch):
    """ASCEND_FFT_VERTICAL: input (N, batch), the transform runs down the columns."""
    x = oracle.uniform_complex((n, batch), seed=n + batch)
    tol = 5e-6 if n == 997 else TOL
    y = c2c(x, [n], FWD, dim_type=S.ASCEND_FFT_VERTICAL)
    ref = oracle.c2c_axis(x, 0, oracle.FORWARD)
    rel, _ = report(f"vertical n={n} b={batch}", y, ref)
    assert rel <= tol


def test_inplace_2d(torch_cuda):
    from fftcase import make, to_dev
    x = oracle.uniform_complex((2, 256, 512), 31)
    h = make([256, 512], S.ASCEND_FFT_C2C, FWD, 2)
    xd = to_dev(x)
    assert h.exec_c2c(xd, xd) == 0
    assert h.synchronize() == 0
    rel, _ = report("in-place 2d", xd.cpu().numpy(), oracle.c2c_nd(x, 2))
    h.destroy()
    assert rel <= TOL


@pytest.mark.parametrize("x", [32, 64, 128])
@pytest.mark.parametrize("y", [32, 64, 128])
def test_fused_2d_tiles(torch_cuda, x, y):
    """Every fused 2D tile instance (both axes of an X x Y slice in one launch), both directions,
    and a 3D transform whose last two axes take the fused path."""
    import sip_amd as S
    from fftcase import make
    b = 5
    a = oracle.uniform_complex((b, x, y), 0x5EED0300 + x * 1000 + y)
    for direction in (FWD, INV):
        h = make([x, y], S.ASCEND_FFT_C2C, direction, b)
        try:
            assert h.describe()["stages"][0].get("tile2d") == [x, y]
        finally:
            h.destroy()
        got = c2c(a, [x, y], direction)
        rel, _ = report(f"fused 2d {x}x{y} b={b} dir={direction:#x}", got, oracle.c2c_nd(a, 2, sign_of(direction)))
        assert rel <= 2e-6
    a3 = oracle.uniform_complex((2, 12, x, y), 0x5EED0400 + x + y)
    got = c2c(a3, [12, x, y], FWD)
    rel, _ = report(f"3d 12x{x}x{y} (fused last two axes)", got, oracle.c2c_nd(a3, 3))
    assert rel <= 2e-6


@pytest.mark.parametrize("split", ["0", "1"])
def test_fused_2d_128_planes(torch_cuda, monkeypatch, split):
    """The 128 x 128 slice with two LDS planes and with one (SIP_FFT_TILE2D_SPLIT), both
    directions, a batch that is not a multiple of anything, and a 3D transform."""
    monkeypatch.setenv("SIP_FFT_TILE2D_SPLIT", split)
    b = 7
    a = oracle.uniform_complex((b, 128, 128), 0x5EED0500 + int(split))
    for direction in (FWD, INV):
        got = c2c(a, [128, 128], direction)
        rel, _ = report(f"fused 2d 128x128 split={split} dir={direction:#x}", got,
                        oracle.c2c_nd(a, 2, sign_of(direction)))
        assert rel <= 2e-6
    a3 = oracle.uniform_complex((2, 6, 128, 128), 0x5EED0600 + int(split))
    got = c2c(a3, [6, 128, 128], INV)
    rel, _ = report(f"3d 6x128x128 split={split}", got, oracle.c2c_nd(a3, 3, sign_of(INV)))
    assert rel <= 2e-6
