"""Serving form (sipFftServe*): continuous batching of FFT requests.

Each request's result is bit-identical whether it runs alone through the server or coalesced with
other requests in the server's mix (same kernels, same per-row arithmetic: long lengths run the
plain plan's kernels, short C2C lengths the persistent multi-length kernel whether alone or mixed),
and within 5e-7 (rel-L2) of the same request through the plain API; a sample is also checked
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


def _served_alone(torch, srv, t, n, b, xd):
    """The request alone through an idle server (nothing else queued or running)."""
    assert srv.drain() == 0
    if t == S.ASCEND_FFT_C2C:
        y = torch.empty((b, n), dtype=torch.complex64, device="cuda")
    else:
        y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
    st, _ = srv.wait(srv.submit(t, n, b, xd.data_ptr(), y.data_ptr()))
    assert st == 0
    return y


def _check_alone(torch, srv, tag, t, n, b, xd, yd):
    """Bit-identical to the request served alone; within 5e-7 of the plain API's result."""
    alone = _served_alone(torch, srv, t, n, b, xd)
    assert torch.equal(alone, yd), f"{tag} (type {t:#x}, n={n}, b={b}) differs alone vs in the mix"
    plain = _alone(torch, t, n, b, xd)
    rel, _ = oracle.error(yd.cpu().numpy(), plain.cpu().numpy().astype(np.complex128))
    assert rel <= 5e-7, f"{tag}: served vs plain API rel-L2 {rel:.3e}"


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
    # one hand-over of the whole burst (the scheduler sees every request at once, so requests of
    # one shape coalesce whatever the host's submission speed)
    reqs = []
    for i, (t, n, b, x, xd, yd) in enumerate(cases):
        reqs.append((t, n, b, xd.data_ptr(), yd.data_ptr()))
        for bt, bn, bb, off in background[3 * i: 3 * i + 3]:
            reqs.append((bt, bn, bb, pool.data_ptr() + 4 * off, None))
    all_tickets = srv.submit_many(reqs)
    tickets = all_tickets[0::4]
    for tk in tickets:
        st, lat = srv.wait(tk)
        assert st == 0 and lat > 0
    assert srv.drain() == 0
    stats = srv.stats()
    assert stats["coalesced_requests"] > stats["launches"], stats   # some launches carried several requests
    for i, (t, n, b, x, xd, yd) in enumerate(cases):
        _check_alone(torch, srv, f"request {i}", t, n, b, xd, yd)
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
    """Requests refused at submit: C2R (not served), n = 0, and a Bluestein length whose workspace
    (157 MB for 20011 x 100: two 65536-point rows per transform) exceeds the 64 MiB budget although
    its input and output would fit, with and without a caller output. The input is a real device
    allocation, so a request that passed validation by mistake would still run in bounds."""
    torch = torch_cuda
    import ctypes
    n = 20011
    x = torch.zeros((100, n), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    srv = S.Server(budget_bytes=1 << 26, nstreams=1, max_coalesce_rows=64)
    t = ctypes.c_uint64()
    L = S.lib()
    xp = ctypes.c_void_p(x.data_ptr())
    assert L.sipFftServeSubmit(srv.h, 0x11, 1024, 1, xp, None, ctypes.byref(t)) != 0  # C2R
    assert L.sipFftServeSubmit(srv.h, 0x10, 0, 1, xp, None, ctypes.byref(t)) != 0
    assert L.sipFftServeSubmit(srv.h, 0x10, n, 100, xp, None, ctypes.byref(t)) != 0   # output + workspace > budget
    assert L.sipFftServeSubmit(srv.h, 0x10, n, 100, xp, ctypes.c_void_p(y.data_ptr()),
                               ctypes.byref(t)) != 0                                 # workspace > budget
    # a batched hand-over with one refused request queues none of them and names the refused one
    with pytest.raises(S.SipFftError, match=r"request 1\)"):
        srv.submit_many([(S.ASCEND_FFT_C2C, 1024, 1, x.data_ptr(), None),
                         (S.ASCEND_FFT_C2C, n, 100, x.data_ptr(), None),
                         (S.ASCEND_FFT_C2C, 1024, 1, x.data_ptr(), None)])
    assert srv.drain() == 0 and srv.stats()["launches"] == 0
    # the same length at a batch whose workspace fits is served and correct
    xs = oracle.uniform_complex((10, n), 31)
    x[:10].copy_(torch.from_numpy(xs))
    tk = srv.submit(S.ASCEND_FFT_C2C, n, 10, x.data_ptr(), y.data_ptr())
    st, _ = srv.wait(tk)
    assert st == 0
    rel, _ = oracle.error(y[:10].cpu().numpy(), oracle.c2c(xs))
    assert rel <= 5e-6, rel
    srv.destroy()


def _watched_inputs(torch, t, n, b, seed):
    if t == S.ASCEND_FFT_C2C:
        x = oracle.uniform_complex((b, n), seed)
        y = torch.empty((b, n), dtype=torch.complex64, device="cuda")
    else:
        x = oracle.uniform(b * n, seed).reshape(b, n)
        y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
    return x, torch.from_numpy(x).cuda(), y


def _check_rows(tag, t, n, x, y):
    """Oracle check of the first and last row of a watched request."""
    rows = sorted({0, x.shape[0] - 1})
    ref = oracle.c2c(x[rows]) if t == S.ASCEND_FFT_C2C else oracle.r2c(x[rows])
    rel, _ = report(tag, y.cpu().numpy()[rows], ref)
    assert rel <= 2e-6


def test_poisson_stream_2k_results(torch_cuda):
    """A seeded Poisson stream at 2,000 requests/s for one second, drawn as the serving configuration
    draws it (70 % C2C / 30 % R2C, n log-uniform over the 2,3,5-smooth lengths in [256, 32768], batch
    U[1, 512], inputs uniform[-1, 1) slices of a device pool). Every 50th request writes to an output
    of its own: each of those is bit-identical to the same request run alone, and its first and last
    rows are within 2e-6 of the FP64 oracle. Every request completes."""
    import time
    torch = torch_cuda
    rng = np.random.Generator(np.random.PCG64(2000))
    sm = S.smooth_lengths(256, 32768)
    pool = torch.rand(256 << 20, device="cuda") * 2 - 1          # 1 GiB of uniform data
    srv = S.Server(budget_bytes=64 << 30, nstreams=4, max_coalesce_rows=8192)
    rate, seconds = 2000.0, 1.0
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        ty = S.ASCEND_FFT_C2C if rng.random() < 0.7 else S.ASCEND_FFT_R2C
        x = np.exp(rng.uniform(np.log(256), np.log(32768)))
        n = next(v for v in sm if v >= x)
        b = int(rng.integers(1, 513))
        reqs.append((t, ty, n, b))
    watched = {}
    for k in range(0, len(reqs), 50):
        _, ty, n, b = reqs[k]
        watched[k] = _watched_inputs(torch, ty, n, b, 7000 + k)
    torch.cuda.synchronize()
    tickets = []
    t0 = time.perf_counter()
    for k, (due, ty, n, b) in enumerate(reqs):
        while time.perf_counter() - t0 < due:
            pass
        if k in watched:
            _, xd, yd = watched[k]
            tickets.append(srv.submit(ty, n, b, xd.data_ptr(), yd.data_ptr()))
        else:
            esz = 8 if ty == S.ASCEND_FFT_C2C else 4
            span = pool.numel() * 4 - n * b * esz
            off = int(rng.integers(0, max(1, span // 8))) * 8
            tickets.append(srv.submit(ty, n, b, pool.data_ptr() + off))
    for tk in tickets:
        st, lat = srv.wait(tk)
        assert st == 0 and lat > 0
    stats = srv.stats()
    print(f"serving 2k/s: {len(reqs)} requests, p50 {stats['p50_ms']:.3f} ms, p99 {stats['p99_ms']:.3f} ms, "
          f"{stats['launches']} launches", flush=True)
    assert stats["completed"] == len(reqs)
    for k, (x, xd, yd) in watched.items():
        _, ty, n, b = reqs[k]
        _check_alone(torch, srv, f"stream request {k}", ty, n, b, xd, yd)
        _check_rows(f"serving 2k request {k} n={n} b={b}", ty, n, x, yd)
    srv.destroy()


def test_admission_waits_for_memory(torch_cuda):
    """A 4 GiB budget and a burst of 28 requests that need 512 MiB each, either of server-owned
    output (C2C 16384 x 4096) or of four-step workspace (C2C 65536 x 1024, caller's output): at most
    eight fit at a time while 16 streams offer 32 launch slots, so memory is what binds and later
    requests wait in the queue until completions free arena space. Every request completes, the
    arena peaks at the budget and never beyond, none is dropped, and the watched results equal the
    requests run alone."""
    torch = torch_cuda
    budget = 4 << 30
    need = 512 << 20
    srv = S.Server(budget_bytes=budget, nstreams=16, max_coalesce_rows=1024)
    pool = torch.rand(128 << 20, device="cuda") * 2 - 1                    # 512 MiB: one C2C 16384 x 4096
    watched, tickets = [], []
    for k in range(4):
        x, xd, yd = _watched_inputs(torch, S.ASCEND_FFT_C2C, 65536, 1024, 8100 + k)
        watched.append((x, xd, yd))
    torch.cuda.synchronize()
    for k in range(24):
        tickets.append(srv.submit(S.ASCEND_FFT_C2C, 16384, 4096, pool.data_ptr()))
        if k % 6 == 5:
            x, xd, yd = watched[k // 6]
            tickets.append(srv.submit(S.ASCEND_FFT_C2C, 65536, 1024, xd.data_ptr(), yd.data_ptr()))
    for tk in tickets:
        st, _ = srv.wait(tk)
        assert st == 0
    stats = srv.stats()
    print(f"admission: peak {stats['arena_peak'] / 2**20:.0f} MiB of {budget / 2**20:.0f}, "
          f"{stats['launches']} launches", flush=True)
    assert stats["completed"] == 28 and stats["launches"] == 28
    assert budget - need < stats["arena_peak"] <= budget, stats
    for k, (x, xd, yd) in enumerate(watched):
        alone = _alone(torch, S.ASCEND_FFT_C2C, 65536, 1024, xd)
        assert torch.equal(alone, yd), f"watched request {k} differs alone vs under admission pressure"
        _check_rows(f"admission watched {k}", S.ASCEND_FFT_C2C, 65536, x, yd)
    srv.destroy()


def test_resident_tables_capped(torch_cuda, monkeypatch):
    """With a zero cap on the resident coefficient tables (only the current launch's tables stay),
    six shapes (four-step, single-pass long rows, Bluestein, R2C) served twice in turn: the other
    shapes' tables are dropped and rebuilt when their shape returns, and every result equals the
    first round's bit for bit and the request run alone through the plain API. (All six shapes'
    tables together take under 1 MiB: the default cap of a sixteenth of the budget, at least
    64 MiB, holds thousands of shapes.)"""
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_SERVE_RESIDENT_MB", "0")
    srv = S.Server(budget_bytes=1 << 30, nstreams=2, max_coalesce_rows=1024)
    shapes = [(S.ASCEND_FFT_C2C, 65536, 4), (S.ASCEND_FFT_C2C, 20000, 300), (S.ASCEND_FFT_C2C, 10007, 4),
              (S.ASCEND_FFT_R2C, 32768, 4), (S.ASCEND_FFT_C2C, 6000, 8), (S.ASCEND_FFT_C2C, 3001, 4)]
    inputs = [_watched_inputs(torch, t, n, b, 8200 + k) for k, (t, n, b) in enumerate(shapes)]
    torch.cuda.synchronize()
    first = []
    for rnd in range(2):
        for k, ((t, n, b), (x, xd, yd)) in enumerate(zip(shapes, inputs)):
            y = torch.empty_like(yd)
            st, _ = srv.wait(srv.submit(t, n, b, xd.data_ptr(), y.data_ptr()))
            assert st == 0
            if rnd == 0:
                first.append(y)
            else:
                assert torch.equal(first[k], y), f"shape {k}: second round differs after eviction"
    stats = srv.stats()
    print(f"resident: {stats['resident_tables']} tables, {stats['resident_bytes'] / 2**20:.2f} MiB, "
          f"{stats['resident_evictions']} evictions", flush=True)
    assert stats["resident_cap"] == 0 and stats["resident_evictions"] > 0
    assert stats["resident_bytes"] < 1 << 20
    for k, ((t, n, b), (x, xd, yd)) in enumerate(zip(shapes, inputs)):
        plain = _alone(torch, t, n, b, xd)
        rel, _ = oracle.error(first[k].cpu().numpy(), plain.cpu().numpy().astype(np.complex128))
        assert rel <= 5e-7, f"shape {k}: served vs plain API rel-L2 {rel:.3e}"
        if n not in (10007, 3001):   # Bluestein lengths are held to the plain API's result above
            _check_rows(f"resident shape {k}", t, n, x, first[k])
    srv.destroy()


def test_coalesced_aux_lengths(torch_cuda):
    """Coalesced launches of lengths whose plans carry auxiliary stages over the gathered rows
    (1, 17 and 997: copy and Bluestein; 20000: four-step): every result equals the request alone."""
    torch = torch_cuda
    srv = S.Server(budget_bytes=4 << 30, nstreams=1, max_coalesce_rows=4096)
    cases = []
    for k, n in enumerate([17, 997, 1, 17, 997, 1, 20000, 20000, 17]):
        x, xd, yd = _watched_inputs(torch, S.ASCEND_FFT_C2C, n, 3 + k % 4, 8300 + k)
        cases.append((n, x, xd, yd))
    torch.cuda.synchronize()
    # a long request first keeps the single stream busy so the rest queue up and coalesce
    big = torch.zeros((64, 65536), dtype=torch.complex64, device="cuda")
    tickets = [srv.submit(S.ASCEND_FFT_C2C, 65536, 64, big.data_ptr())]
    for n, x, xd, yd in cases:
        tickets.append(srv.submit(S.ASCEND_FFT_C2C, n, x.shape[0], xd.data_ptr(), yd.data_ptr()))
    for tk in tickets:
        st, _ = srv.wait(tk)
        assert st == 0
    for n, x, xd, yd in cases:
        alone = _alone(torch, S.ASCEND_FFT_C2C, n, x.shape[0], xd)
        assert torch.equal(alone, yd), f"n={n}: coalesced result differs from the request alone"
        rel, _ = report(f"coalesced aux n={n}", yd.cpu().numpy(), oracle.c2c(x))
        assert rel <= (5e-6 if n in (17, 997) else 2e-6)
    srv.destroy()


def test_mixed_length_launches(torch_cuda, monkeypatch):
    """With SIP_FFT_SERVE_MIXED=1, short C2C requests of different lengths (2,3,5-smooth up to 4096)
    queued behind a long one coalesce into persistent multi-length launches: fewer launches than
    requests, at least one mixed launch, and every result bit-identical to the request served
    alone, within 5e-7 of the plain API and within 2e-6 of the FP64 oracle."""
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_SERVE_MIXED", "1")
    srv = S.Server(budget_bytes=4 << 30, nstreams=1, max_coalesce_rows=4096)
    lengths = [256, 1536, 4096, 100, 3000, 243, 2000, 625, 4096, 256, 1000, 8, 720, 1024, 30, 64]
    cases = []
    for k, n in enumerate(lengths):
        x, xd, yd = _watched_inputs(torch, S.ASCEND_FFT_C2C, n, 1 + (k * 7) % 13, 8500 + k)
        cases.append((n, x, xd, yd))
    torch.cuda.synchronize()
    big = torch.zeros((64, 65536), dtype=torch.complex64, device="cuda")
    tickets = [srv.submit(S.ASCEND_FFT_C2C, 65536, 64, big.data_ptr())]     # keeps the stream busy
    for n, x, xd, yd in cases:
        tickets.append(srv.submit(S.ASCEND_FFT_C2C, n, x.shape[0], xd.data_ptr(), yd.data_ptr()))
    for tk in tickets:
        st, _ = srv.wait(tk)
        assert st == 0
    stats = srv.stats()
    print(f"mixed: {stats['launches']} launches ({stats['mixed_launches']} mixed) for {len(tickets)} requests",
          flush=True)
    assert stats["mixed_launches"] >= 1 and stats["launches"] < len(tickets), stats
    for n, x, xd, yd in cases:
        _check_alone(torch, srv, f"mixed n={n}", S.ASCEND_FFT_C2C, n, x.shape[0], xd, yd)
        rel, _ = report(f"mixed n={n}", yd.cpu().numpy(), oracle.c2c(x))
        assert rel <= 2e-6
    srv.destroy()
