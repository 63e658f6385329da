"""Device-resident coefficient cache (the counterpart of the reference's FFTensorCache,
core/fft/utils/fftensor_cache.cpp:38-92): two live plans of one shape share one device table, the
table outlives the first plan's destruction and is freed with the last plan, and the map drops the
expired entry when the next table is created. Results through the shared table are checked against
the FP64 oracle (rel-L2 <= 2e-6).
"""
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, make, report

pytestmark = pytest.mark.gpu


def _run(torch, h, n, b, seed):
    x = oracle.uniform_complex((b, n), seed)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    assert h.exec_c2c(xd, y) == S.ACL_SUCCESS
    assert h.synchronize() == S.ACL_SUCCESS
    rel, _ = report(f"cache n={n} b={b}", y.cpu().numpy(), oracle.c2c(x))
    assert rel <= 2e-6


def test_tables_shared_and_released_with_last_plan(torch_cuda):
    torch = torch_cuda
    n = 1875  # leaf 15-5-5-5: one pass-twiddle table, used by no other test
    _, live0, bytes0 = S.coeff_cache_stats()
    h1 = make([n], S.ASCEND_FFT_C2C, FWD, 4)
    _, live1, bytes1 = S.coeff_cache_stats()
    assert live1 == live0 + 1 and bytes1 > bytes0
    h2 = make([n], S.ASCEND_FFT_C2C, FWD, 9)      # another batch: the same table, nothing uploaded
    _, live2, bytes2 = S.coeff_cache_stats()
    assert (live2, bytes2) == (live1, bytes1)
    _run(torch, h1, n, 4, 0x5EED0C01)
    _run(torch, h2, n, 9, 0x5EED0C02)
    h1.destroy()                                  # h2 still holds the table
    _, live3, bytes3 = S.coeff_cache_stats()
    assert (live3, bytes3) == (live1, bytes1)
    _run(torch, h2, n, 9, 0x5EED0C03)
    h2.destroy()                                  # last user: freed
    entries4, live4, bytes4 = S.coeff_cache_stats()
    assert (live4, bytes4) == (live0, bytes0)
    # the next table created drops the expired entry from the map
    h3 = make([1536], S.ASCEND_FFT_C2C, FWD, 2)
    entries5, _, _ = S.coeff_cache_stats()
    assert entries5 <= entries4
    h3.destroy()


def test_inverse_uses_its_own_table(torch_cuda):
    """Forward and inverse plans of one length hold different tables (conjugate twiddles)."""
    n = 1875
    _, live0, _ = S.coeff_cache_stats()
    hf = make([n], S.ASCEND_FFT_C2C, FWD, 2)
    hi = make([n], S.ASCEND_FFT_C2C, S.ASCEND_FFT_INVERSE, 2)
    _, live1, _ = S.coeff_cache_stats()
    assert live1 == live0 + 2
    hf.destroy()
    hi.destroy()
    _, live2, _ = S.coeff_cache_stats()
    assert live2 == live0
