"""GPU parity of the plain form (Stockham kernels through the C-ABI) against the FP64 oracle.

Tolerances (rel-L2 vs FP64): 1e-6 for C2C N=1024 b=256, 2e-6 for every other plain transform,
as stated for this path; Bluestein lengths (prime factors > 13) 5e-6. Round trips
inverse(forward(x)) / N are checked against x at 2e-6.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, c2c, c2r, make, r2c, report, sign_of, to_dev

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _c2c_case(n, batch, direction, seed, tol=TOL):
    x = oracle.uniform_complex((batch, n), seed)
    y = c2c(x, [n], direction)
    ref = oracle.c2c(x, sign_of(direction))
    rel, _ = report(f"c2c n={n} b={batch} dir={direction:#x}", y, ref)
    assert rel <= tol
    return x, y


@pytest.mark.parametrize("direction", [FWD, INV])
def test_c1_c2c_1024_b256(torch_cuda, direction):
    _c2c_case(1024, 256, direction, seed=0x5EED0001, tol=1e-6)


def test_c2_c2c_4096_b16384_full(torch_cuda):
    """Headline workload at full size: every output bin against the FP64 oracle, and the round trip."""
    n, b = 4096, 16384
    x, y = _c2c_case(n, b, FWD, seed=0x5EED0002)
    back = c2c(y, [n], INV) / n
    rel, _ = report("roundtrip 4096x16384", back, x.astype(np.complex128))
    assert rel <= TOL


@pytest.mark.parametrize("n", [1536, 2000])
def test_c2_mixed_radix_b8192(torch_cuda, n):
    x, y = _c2c_case(n, 8192, FWD, seed=0x5EED0003 + n)
    back = c2c(y, [n], INV) / n
    rel, _ = report(f"roundtrip {n}x8192", back, x.astype(np.complex128))
    assert rel <= TOL


def test_c2_r2c_c2r_8192_b4096(torch_cuda):
    n, b = 8192, 4096
    x = oracle.uniform(n * b, 0x5EED0004).reshape(b, n)
    X = r2c(x, [n])
    rel, _ = report("r2c 8192x4096", X, oracle.r2c(x))
    assert rel <= TOL
    back = c2r(X, [n]) / n
    rel, _ = report("c2r roundtrip 8192x4096", back, x.astype(np.float64))
    assert rel <= TOL
    # C2R against the oracle on the GPU's own R2C output
    y = c2r(X, [n])
    rel, _ = report("c2r 8192x4096", y, oracle.c2r(X, n))
    assert rel <= TOL


POW2 = [2 ** k for k in range(1, 15)]
SMOOTH = [3, 5, 6, 9, 10, 12, 15, 25, 27, 45, 60, 81, 100, 125, 243, 360, 625, 768, 1000, 1080, 2187, 2400,
          3000, 3125, 3840, 4000, 4374, 5120, 6000, 9000, 12288, 13122, 15360, 16000]
PRIME_FACTOR = [7, 11, 13, 14, 21, 49, 77, 91, 143, 169, 343, 1001, 1331, 2197, 4004]


@pytest.mark.parametrize("n", POW2 + SMOOTH + PRIME_FACTOR)
def test_c2c_leaf_lengths(torch_cuda, n):
    batch = max(1, min(64, (1 << 18) // n))
    for direction in (FWD, INV):
        _c2c_case(n, batch, direction, seed=1000 + n)


@pytest.mark.parametrize("n", [1, 17, 31, 97, 127, 997, 199 * 199, 6007])
def test_c2c_bluestein_and_trivial(torch_cuda, n):
    batch = 3
    _c2c_case(n, batch, FWD, seed=2000 + n, tol=5e-6)
    _c2c_case(n, batch, INV, seed=3000 + n, tol=5e-6)


@pytest.mark.parametrize("n", [32000, 65536, 3 ** 10, 5 ** 7, 2 ** 20, 15000, 2 ** 17 * 3])
def test_c2c_four_step(torch_cuda, n):
    batch = max(1, (1 << 21) // n)
    x, y = _c2c_case(n, batch, FWD, seed=4000 + n)
    back = c2c(y, [n], INV) / n
    rel, _ = report(f"roundtrip {n}", back, x.astype(np.complex128))
    assert rel <= TOL


@pytest.mark.parametrize("n", [16875, 19683, 20000, 24576, 30375, 32768])
def test_c2c_single_pass_long_rows(torch_cuda, n):
    """Single-launch rows above 16384 (one 1024-thread workgroup per transform, one LDS plane, from
    256 transforms up): forward and inverse against the oracle, the round trip, and the same length
    as the last axis of a 2D transform."""
    b = 257
    x, y = _c2c_case(n, b, FWD, seed=4500 + n)
    _c2c_case(n, b, INV, seed=4600 + n)
    back = c2c(y, [n], INV) / n
    rel, _ = report(f"roundtrip {n}", back, x.astype(np.complex128))
    assert rel <= TOL
    x2 = oracle.uniform_complex((86, 3, n), 4700 + n)
    y2 = c2c(x2, [3, n], FWD)
    rel, _ = report(f"2D 3x{n}", y2, oracle.c2c_nd(x2, 2, oracle.FORWARD))
    assert rel <= TOL


def _measured_lengths():
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sip_amd", "csrc", "kernels",
                        "gen_instances.py")
    spec = importlib.util.spec_from_file_location("gen_instances_for_test", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return sorted(mod.SEQ_OVERRIDE)


@pytest.mark.parametrize("n", _measured_lengths())
def test_c2c_measured_radix_sequences(torch_cuda, n):
    """Lengths whose radix sequence was chosen by measurement: forward against the oracle and the
    round trip, at a batch that runs them as leaves (256 rows from 16384 up)."""
    b = 256 if n > 16384 else max(2, min(64, (1 << 20) // n))
    x, y = _c2c_case(n, b, FWD, seed=8800 + n)
    back = c2c(y, [n], INV) / n
    rel, _ = report(f"roundtrip {n}", back, x.astype(np.complex128))
    assert rel <= TOL


@pytest.mark.parametrize("dims,b", [([65536], 2048), ([1024, 1024], 128), ([512, 512], 256), ([128, 128, 128], 64)])
def test_streaming_column_outputs_1gib(torch_cuda, monkeypatch, dims, b):
    """1 GiB C2C plans whose last stage stores column tiles nontemporally: bit-equal to the same
    transform with cached stores (SIP_FFT_NT_OUT=0 at plan time), and three batch entries against
    the oracle."""
    torch = torch_cuda
    x = oracle.uniform_complex((b, *dims), 0x5EED0A00 + len(dims))
    xd = to_dev(x)
    outs = []
    for nt in ("1", "0"):
        monkeypatch.setenv("SIP_FFT_NT_OUT", nt)
        h = make(dims, S.ASCEND_FFT_C2C, FWD, b)
        try:
            stages = h.describe()["stages"]
            assert stages[-1]["nt_out"] == (1 if nt == "1" else 0) and len(stages) >= 2
            yd = torch.empty_like(xd)
            assert h.exec_c2c(xd, yd) == S.ACL_SUCCESS and h.synchronize() == S.ACL_SUCCESS
            outs.append(yd)
        finally:
            h.destroy()
    assert torch.equal(outs[0], outs[1]), "nontemporal stores changed the result"
    for k in (0, b // 2, b - 1):
        got = outs[0][k].cpu().numpy()
        ref = oracle.c2c_nd(x[k], len(dims), oracle.FORWARD)
        rel, _ = report(f"c2c {dims} x {b} entry {k}", got, ref)
        assert rel <= TOL


def test_c2c_2_pow_27(torch_cuda):
    """Largest size the reference accepts (0 < N <= 2^27): three-level four-step."""
    n = 1 << 27
    x = oracle.uniform_complex((1, n), 77)
    y = c2c(x, [n], FWD)
    ref = oracle.c2c(x, oracle.FORWARD)
    rel, _ = report("c2c 2^27", y, ref)
    assert rel <= TOL


@pytest.mark.parametrize("n", [4, 6, 8, 10, 16, 126, 128, 256, 500, 1024, 2000, 4096, 8192, 16384, 18432, 20000,
                               24000, 24576, 32768, 65536, 3, 5, 7, 9, 15, 243, 567, 997, 2187, 199 * 199, 15000])
def test_r2c_c2r_lengths(torch_cuda, n):
    b = max(1, min(32, (1 << 19) // n))
    x = oracle.uniform(n * b, 5000 + n).reshape(b, n)
    tol = 5e-6 if n in (997, 199 * 199) else TOL
    X = r2c(x, [n])
    rel, _ = report(f"r2c n={n}", X, oracle.r2c(x))
    assert rel <= tol
    Xh = oracle.r2c(x).astype(np.complex64)
    y = c2r(Xh, [n])
    rel, _ = report(f"c2r n={n}", y, oracle.c2r(Xh, n))
    assert rel <= tol


@pytest.mark.parametrize("n,b", [(1024, 7), (2048, 5), (16384, 3), (8192, 1)])
def test_r2c_c2r_ragged_batch(torch_cuda, n, b):
    """Batches that leave the last tile of several transforms partly empty."""
    x = oracle.uniform(n * b, 7000 + n + b).reshape(b, n)
    X = r2c(x, [n])
    rel, _ = report(f"r2c n={n} b={b}", X, oracle.r2c(x))
    assert rel <= TOL
    Xh = oracle.r2c(x).astype(np.complex64)
    y = c2r(Xh, [n])
    rel, _ = report(f"c2r n={n} b={b}", y, oracle.c2r(Xh, n))
    assert rel <= TOL


def test_reference_example_input(torch_cuda):
    """The reference example's input x[i] = (i, i+1) over batch 32 x N 128
    (example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:99-110), against the committed golden vectors."""
    from golden import load_fixture
    fx = load_fixture("example_c2c_128x32")
    y = c2c(fx["input"], [128], FWD)
    rel, _ = report("example c2c 128x32", y, fx["output"])
    assert rel <= 1e-6
    fx = load_fixture("example_r2c_256x32")
    y = r2c(fx["input"], [256], FWD)
    rel, _ = report("example r2c 256x32", y, fx["output"])
    assert rel <= 1e-6
    fx = load_fixture("example_c2r_128x32_forward")
    y = c2r(fx["input"], [128], FWD)
    rel, _ = report("example c2r 128x32 forward", y, fx["output"])
    assert rel <= 1e-6


@pytest.mark.parametrize("n", [1000, 4096])
def test_direction_quirks(torch_cuda, n):
    """R2C with INVERSE computes the e^{+} transform; C2R with FORWARD the e^{-} transform of the
    Hermitian extension (dft_r2c_core.cpp:96-101, dft_c2r_core.cpp:90-105)."""
    b = 4
    x = oracle.uniform(n * b, 99).reshape(b, n)
    X = r2c(x, [n], INV)
    rel, _ = report("r2c inverse", X, oracle.r2c(x, oracle.INVERSE))
    assert rel <= TOL
    Xh = oracle.r2c(x).astype(np.complex64)
    y = c2r(Xh, [n], FWD)
    rel, _ = report("c2r forward", y, oracle.c2r(Xh, n, oracle.FORWARD))
    assert rel <= TOL


def test_inplace_c2c(torch_cuda):
    import sip_amd as S
    from fftcase import make, to_dev
    n, b = 4096, 64
    x = oracle.uniform_complex((b, n), 4242)
    h = make([n], S.ASCEND_FFT_C2C, FWD, b)
    xd = to_dev(x)
    assert h.exec_c2c(xd, xd) == 0
    assert h.synchronize() == 0
    rel, _ = report("in-place 4096", xd.cpu().numpy(), oracle.c2c(x))
    h.destroy()
    assert rel <= TOL


def test_unaligned_rows(torch_cuda):
    """Rows 8-byte but not 16-byte aligned (views one element into a buffer): the direct passes
    fall back to element-wise HBM access and give the same result as aligned rows."""
    import torch
    import sip_amd as S
    from fftcase import make
    n, b = 4096, 32
    x = oracle.uniform_complex((b, n), 4343)
    buf_in = torch.zeros(b * n + 1, dtype=torch.complex64, device="cuda")
    buf_out = torch.zeros(b * n + 1, dtype=torch.complex64, device="cuda")
    xin = buf_in[1:].view(b, n)
    xin.copy_(torch.from_numpy(x).cuda())
    yout = buf_out[1:].view(b, n)
    assert xin.data_ptr() % 16 == 8 and yout.data_ptr() % 16 == 8
    h = make([n], S.ASCEND_FFT_C2C, FWD, b)
    try:
        assert h.exec_c2c(xin, yout) == 0
        assert h.synchronize() == 0
        y_unaligned = yout.cpu().numpy()
        xa = xin.contiguous().clone()
        ya = torch.empty_like(xa)
        assert h.exec_c2c(xa, ya) == 0
        assert h.synchronize() == 0
    finally:
        h.destroy()
    rel, _ = report("unaligned rows 4096x32", y_unaligned, oracle.c2c(x))
    assert rel <= TOL
    assert np.array_equal(y_unaligned, ya.cpu().numpy())
