"""Serving form (sipFftServe*): continuous batching of FFT requests.

Each request's result is bit-identical whether it runs alone (plain plan) or coalesced with other
requests in the server's mix (same kernels, same per-row arithmetic); a sample is also checked
against the FP64 oracle (rel-L2 <= 2e-6). The open-loop benchmark completes every request it admits
at a modest rate, and latency percentiles are ordered.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, report

pytestmark = pytest.mark.gpu


def _alone(torch, t, n, b, x):
    h = S.FftHandle()
    assert h.make_plan_1d(n, t, FWD, b) == 0
    h.alloc_workspace()
    h.use_current_torch_stream()
    if t == S.ASCEND_FFT_C2C:
        y = torch.empty((b, n), dtype=torch.complex64, device="cuda")
        assert h.exec_c2c(x, y) == 0
    else:
        y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
        assert h.exec_r2c(x, y) == 0
    assert h.synchronize() == 0
    h.destroy()
    return y


def test_alone_equals_mix(torch_cuda):
    torch = torch_cuda
    rng = np.random.Generator(np.random.PCG64(11))
    sm = S.smooth_lengths(256, 32768)
    srv = S.Server(budget_bytes=16 << 30, nstreams=4, max_coalesce_rows=4096)
    pool = torch.rand(64 << 20, device="cuda") * 2 - 1          # 256 MiB of uniform data
    cases = []
    # watched requests: explicit outputs, repeated shapes so that coalescing happens
    shapes = [(S.ASCEND_FFT_C2C, 4096), (S.ASCEND_FFT_R2C, 8192), (S.ASCEND_FFT_C2C, 1536), (S.ASCEND_FFT_R2C, 375),
              (S.ASCEND_FFT_C2C, 30000), (S.ASCEND_FFT_C2C, 256), (S.ASCEND_FFT_R2C, 32768)]
    # prepare every input first, then submit one burst so that requests queue up and coalesce
    background = []
    for i in range(70):
        t, n = shapes[i % len(shapes)]
        b = int(rng.integers(1, 24))
        if t == S.ASCEND_FFT_C2C:
            x = oracle.uniform_complex((b, n), 900 + i)
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty((b, n), dtype=torch.complex64, device="cuda")
        else:
            x = oracle.uniform(b * n, 900 + i).reshape(b, n)
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
        cases.append((t, n, b, x, xd, yd))
        for _ in range(3):
            bt = S.ASCEND_FFT_C2C if rng.random() < 0.7 else S.ASCEND_FFT_R2C
            bn = int(sm[int(rng.integers(0, len(sm)))])
            bb = int(rng.integers(1, 64))
            off = int(rng.integers(0, pool.numel() - 2 * bn * bb - 2)) // 2 * 2
            background.append((bt, bn, bb, off))
    torch.cuda.synchronize()
    tickets = []
    for i, (t, n, b, x, xd, yd) in enumerate(cases):
        tickets.append(srv.submit(t, n, b, xd.data_ptr(), yd.data_ptr()))
        for bt, bn, bb, off in background[3 * i: 3 * i + 3]:
            srv.submit(bt, bn, bb, pool.data_ptr() + 4 * off)
    for tk in tickets:
        st, lat = srv.wait(tk)
        assert st == 0 and lat > 0
    assert srv.drain() == 0
    stats = srv.stats()
    assert stats["coalesced_requests"] > stats["launches"], stats   # some launches carried several requests
    for i, (t, n, b, x, xd, yd) in enumerate(cases):
        alone = _alone(torch, t, n, b, xd)
        assert torch.equal(alone, yd), f"request {i} (type {t:#x}, n={n}, b={b}) differs alone vs in the mix"
        if i < 14:
            ref = oracle.c2c(x) if t == S.ASCEND_FFT_C2C else oracle.r2c(x)
            rel, _ = report(f"serving request {i} n={n}", yd.cpu().numpy(), ref)
            assert rel <= 2e-6
    srv.destroy()


def test_benchmark_modest_rate(torch_cuda):
    torch = torch_cuda
    srv = S.Server(budget_bytes=16 << 30, nstreams=4, max_coalesce_rows=8192)
    pool = torch.rand(128 << 20, device="cuda") * 2 - 1
    torch.cuda.synchronize()
    r = srv.benchmark(rate=50.0, seconds=1.0, pool_ptr=pool.data_ptr(), pool_bytes=pool.numel() * 4, seed=3,
                      drain_seconds=20.0)
    assert r["status"] == 0 and r["dropped_unadmitted"] == 0 and r["completed"] == r["submitted"] > 10, r
    sv = r["server"]
    assert 0 < sv["p50_ms"] <= sv["p90_ms"] <= sv["p99_ms"] <= sv["max_ms"]
    srv.destroy()


def test_submit_validation(torch_cuda):
    srv = S.Server(budget_bytes=1 << 26, nstreams=1, max_coalesce_rows=64)
    import ctypes
    t = ctypes.c_uint64()
    L = S.lib()
    assert L.sipFftServeSubmit(srv.h, 0x11, 1024, 1, ctypes.c_void_p(0x1000), None, ctypes.byref(t)) != 0  # C2R
    assert L.sipFftServeSubmit(srv.h, 0x10, 0, 1, ctypes.c_void_p(0x1000), None, ctypes.byref(t)) != 0
    assert L.sipFftServeSubmit(srv.h, 0x10, 1 << 20, 512, ctypes.c_void_p(0x1000), None, ctypes.byref(t)) != 0  # > budget
    srv.destroy()
