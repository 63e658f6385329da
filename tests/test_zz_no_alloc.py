"""The no-allocation C2C entry point (sipFftExecC2CNoAlloc) through the C-ABI, against
asdFftExecC2C with a caller workspace on the same plan shape.

Both run the same kernels on the same input, so the outputs must be bit-identical; each is also
checked against the FP64 oracle (rel-L2 <= 2e-6, Bluestein lengths <= 5e-6). Two calls in a row per
shape kind, the second with a smaller input, show that the shared scratch array (which keeps the
previous call's contents) does not leak into a result. A plan whose workspace exceeds the 64 MiB
scratch is refused before anything runs.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, make, report, sign_of

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _pair(torch, n, b, direction, seed, tol):
    print(f"[no-alloc] n={n} b={b} dir={direction:#x}", flush=True)
    x = oracle.uniform_complex((b, n), seed)
    xd = torch.from_numpy(x).cuda()
    y_ref = torch.empty_like(xd)
    y_got = torch.full_like(xd, float("nan"))
    ref_h = make([n], S.ASCEND_FFT_C2C, direction, b)            # caller workspace bound
    assert 0 < ref_h.workspace_size() <= SCRATCH
    h = S.FftHandle()                                           # no workspace bound at all
    assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, direction, b) == S.ACL_SUCCESS
    h.use_current_torch_stream()
    assert h.exec_c2c(xd, y_got) == S.ACL_ERROR_INTERNAL_ERROR  # the plain entry needs the workspace
    assert ref_h.exec_c2c(xd, y_ref) == S.ACL_SUCCESS
    assert h.exec_c2c_noalloc(xd, y_got) == S.ACL_SUCCESS
    torch.cuda.synchronize()
    a, g = y_ref.cpu().numpy(), y_got.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), g.view(np.uint32)), "no-alloc result differs from asdFftExecC2C"
    rel, _ = report(f"no-alloc n={n} b={b}", g, oracle.c2c(x, sign_of(direction)))
    assert rel <= tol
    assert torch.equal(xd.cpu(), torch.from_numpy(x))          # the input is not modified
    ref_h.destroy()
    h.destroy()


@pytest.mark.parametrize("big,small,tol", [((65536, 16), (3 ** 10, 8), 2e-6),   # four-step intermediate
                                           ((997, 64), (997, 8), 5e-6),        # Bluestein rows
                                           ((6007, 40), (5003, 8), 5e-6)])
def test_no_alloc_matches_workspace_exec(torch_cuda, big, small, tol):
    torch = torch_cuda
    _pair(torch, big[0], big[1], FWD, 0x5EED0A01, tol)
    _pair(torch, small[0], small[1], INV, 0x5EED0A02, tol)


def test_no_alloc_without_workspace_need(torch_cuda):
    """A single-launch plan (no workspace) runs through the entry point unchanged."""
    torch = torch_cuda
    n, b = 4096, 32
    x = oracle.uniform_complex((b, n), 0x5EED0A03)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    h = S.FftHandle()
    assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, FWD, b) == S.ACL_SUCCESS
    assert h.workspace_size() == 0
    h.use_current_torch_stream()
    assert h.exec_c2c_noalloc(xd, y) == S.ACL_SUCCESS
    rel, _ = report("no-alloc 4096x32", y.cpu().numpy(), oracle.c2c(x))
    assert rel <= 2e-6
    h.destroy()


def test_no_alloc_refuses_beyond_scratch(torch_cuda):
    """65536 x 256 needs a 128 MiB four-step intermediate: refused with 100000, output untouched."""
    torch = torch_cuda
    n, b = 65536, 256
    h = S.FftHandle()
    assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, FWD, b) == S.ACL_SUCCESS
    assert h.workspace_size() > SCRATCH
    h.use_current_torch_stream()
    xd = torch.zeros((b, n), dtype=torch.complex64, device="cuda")
    y = torch.full_like(xd, 7.0)
    assert h.exec_c2c_noalloc(xd, y) == S.ACL_ERROR_INVALID_PARAM
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())
    h.destroy()


def test_no_alloc_validation(torch_cuda):
    """Same argument checks as asdFftExecC2C: bad handle, wrong plan type, wrong shape."""
    torch = torch_cuda
    xd = torch.zeros((8, 1024), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(xd)
    h = S.FftHandle()
    assert h.make_plan_1d(1024, S.ASCEND_FFT_R2C, FWD, 8) == S.ACL_SUCCESS
    assert h.exec_c2c_noalloc(xd, y) == S.ACL_ERROR_INVALID_PARAM       # R2C plan
    assert h.make_plan_1d(2048, S.ASCEND_FFT_C2C, FWD, 8) == S.ACL_SUCCESS
    assert h.exec_c2c_noalloc(xd, y) == S.ACL_ERROR_INVALID_PARAM       # shape mismatch
    h.destroy()
    assert S.lib().sipFftExecC2CNoAlloc(None, None, None) == S.ACL_ERROR_INVALID_PARAM
