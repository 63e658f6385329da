"""Batching form (sipFftBatch*): uneven C2C problems in one launch.

Checks: every problem against the FP64 oracle (rel-L2 <= 2e-6); the batched result is bit-identical
to each problem run alone through asdFftExecC2C when both take the same plan (same kernels, same
per-transform arithmetic; the problems here keep every group of lengths 16384..32768 below the 64
rows from which such a group runs as a single-pass leaf in a batch instead of a four-step split,
256 in a plain plan);
hipGraph replays (with eager work and synchronisation in between) reproduce the eager result bit
for bit; autotuned plans stay within tolerance and reuse the on-disk cache.
"""
import numpy as np
import pytest

import oracle
import sip_amd as S
from fftcase import FWD, INV, report

pytestmark = pytest.mark.gpu
TOL = 2e-6


def _problems(torch, lengths, batches, seed):
    xs, ys, hs = [], [], []
    for i, (n, b) in enumerate(zip(lengths, batches)):
        x = oracle.uniform_complex((b, n), seed + i)
        hs.append(x)
        xs.append(torch.from_numpy(x).cuda())
        ys.append(torch.empty((b, n), dtype=torch.complex64, device="cuda"))
    return hs, xs, ys


def _run_batch(torch, lengths, batches, xs, ys, direction, flags):
    bh = S.BatchHandle()
    st = bh.make_plan(lengths, batches, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys], direction, flags)
    assert st == 0, st
    bh.alloc_workspace()
    s = torch.cuda.current_stream().cuda_stream
    if s:
        assert bh.set_stream(s) == 0
    assert bh.exec() == 0
    torch.cuda.synchronize()
    return bh


LENGTHS = [64, 100, 4096, 4096, 1536, 2000, 8192, 16384, 20000, 65536, 3000, 128, 64, 59049 // 243 * 243, 25,
           30000, 45, 1024, 1000, 36]
BATCHES = [3, 1, 2, 7, 5, 4, 2, 1, 2, 1, 6, 9, 2, 1, 8, 1, 3, 4, 2, 5]


@pytest.mark.parametrize("direction", [FWD, INV])
def test_batch_matches_oracle_and_alone(torch_cuda, direction):
    torch = torch_cuda
    hs, xs, ys = _problems(torch, LENGTHS, BATCHES, 100)
    bh = _run_batch(torch, LENGTHS, BATCHES, xs, ys, direction, S.BATCH_GRAPH)
    d = bh.describe()
    assert d["groups"] == len(set(LENGTHS)) and d["graph"] == 1
    for i, (n, b) in enumerate(zip(LENGTHS, BATCHES)):
        got = ys[i].cpu().numpy()
        rel, _ = report(f"batch problem {i} n={n} b={b}", got, oracle.c2c(hs[i], oracle.FORWARD if direction == FWD
                                                                          else oracle.INVERSE))
        assert rel <= TOL
        # alone: the same problem through a plain plan must give the same bits
        h = S.FftHandle()
        assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, direction, b) == 0
        h.alloc_workspace()
        h.use_current_torch_stream()
        alone = torch.empty_like(ys[i])
        assert h.exec_c2c(xs[i], alone) == 0
        assert h.synchronize() == 0
        assert torch.equal(alone, ys[i]), f"problem {i}: batched result differs from the problem alone"
        h.destroy()
    bh.destroy()


def test_graph_replays_match_eager(torch_cuda):
    torch = torch_cuda
    hs, xs, ys = _problems(torch, LENGTHS, BATCHES, 300)
    eager_ys = [torch.empty_like(y) for y in ys]
    be = _run_batch(torch, LENGTHS, BATCHES, xs, eager_ys, FWD, 0)          # direct launches
    bg = _run_batch(torch, LENGTHS, BATCHES, xs, ys, FWD, S.BATCH_GRAPH)    # graph
    for rep in range(4):
        for y in ys:
            y.zero_()
        junk = torch.randn(1 << 20, device="cuda") * 3  # eager work between replays
        torch.cuda.synchronize()
        assert bg.exec() == 0
        torch.cuda.synchronize()
        for i in range(len(ys)):
            assert torch.equal(ys[i], eager_ys[i]), f"replay {rep} problem {i}"
        del junk
    be.destroy()
    bg.destroy()


def test_autotune_and_cache(torch_cuda, tmp_path, monkeypatch):
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    lengths = [4096, 1024, 65536, 1536, 20000]
    batches = [64, 256, 4, 32, 3]
    hs, xs, ys = _problems(torch, lengths, batches, 500)
    bh = _run_batch(torch, lengths, batches, xs, ys, FWD, S.BATCH_GRAPH | S.BATCH_AUTOTUNE)
    d = bh.describe()
    assert all(g["tuned_us"] > 0 for g in d["group_list"]), d
    for i in range(len(lengths)):
        rel, _ = report(f"tuned problem {i}", ys[i].cpu().numpy(), oracle.c2c(hs[i]))
        assert rel <= TOL
    bh.destroy()
    lines = (tmp_path / "tune.jsonl").read_text().strip().splitlines()
    assert len(lines) == len(lengths)


def test_batching_workload_shape():
    lengths, batches = S.batching_workload(4096)
    sm = set(S.smooth_lengths(64, 65536))
    assert len(lengths) == 4096 and all(n in sm for n in lengths) and all(1 <= b <= 64 for b in batches)
    assert lengths == S.batching_workload(4096)[0]


def test_autotune_leaves_no_pending_device_error(torch_cuda, tmp_path, monkeypatch):
    """Tuning tries tile shapes that the device may refuse (too much LDS, too many threads); the
    refusals are skipped and must not surface in the caller's next HIP call."""
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    lengths = [8192, 16384, 12000, 15000, 6000, 3000, 2400, 1080, 65536, 43200]
    batches = [4, 2, 3, 2, 5, 6, 8, 9, 1, 1]
    hs, xs, ys = _problems(torch, lengths, batches, 900)
    bh = _run_batch(torch, lengths, batches, xs, ys, FWD, S.BATCH_GRAPH | S.BATCH_AUTOTUNE)
    torch.zeros(1, device="cuda").add_(1)   # raises if a HIP error is pending
    torch.cuda.synchronize()
    for i in range(len(lengths)):
        rel, _ = report(f"tuned problem {lengths[i]}", ys[i].cpu().numpy(), oracle.c2c(hs[i]))
        assert rel <= TOL
    bh.destroy()


def test_gathered_aux_stages(torch_cuda):
    """Lengths whose plans carry auxiliary stages over the gathered rows: 1 (a copy), 17, 997 and 6007
    (Bluestein chirp, kernel and post stages) and 20000 (four-step); problems interleaved in memory
    so that a stage reading or writing the wrong row shows up."""
    torch = torch_cuda
    lengths = [17, 1, 997, 17, 1, 6007, 997, 20000, 17]
    batches = [3, 5, 2, 4, 2, 1, 3, 2, 1]
    hs, xs, ys = _problems(torch, lengths, batches, 1300)
    for direction in (FWD, INV):
        for y in ys:
            y.fill_(float("nan"))
        bh = _run_batch(torch, lengths, batches, xs, ys, direction, S.BATCH_GRAPH)
        sgn = oracle.FORWARD if direction == FWD else oracle.INVERSE
        for i, (n, b) in enumerate(zip(lengths, batches)):
            rel, _ = report(f"gathered aux n={n} b={b} dir={direction:#x}", ys[i].cpu().numpy(),
                            oracle.c2c(hs[i], sgn))
            assert rel <= (5e-6 if n in (17, 997, 6007) else TOL)
        bh.destroy()


def test_autotune_in_place(torch_cuda, tmp_path, monkeypatch):
    """In-place problems (output buffer = input buffer) with autotuning: the tuner times its
    candidates on scratch rows, so the caller's data is transformed exactly once, by the exec."""
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    lengths = [4096, 1000, 65536, 997, 1536]
    batches = [8, 6, 2, 3, 5]
    hs, xs, _ = _problems(torch, lengths, batches, 1500)
    bh = _run_batch(torch, lengths, batches, xs, xs, FWD, S.BATCH_GRAPH | S.BATCH_AUTOTUNE)
    for i, (n, b) in enumerate(zip(lengths, batches)):
        rel, _ = report(f"in-place tuned n={n}", xs[i].cpu().numpy(), oracle.c2c(hs[i]))
        assert rel <= (5e-6 if n == 997 else TOL)
    bh.destroy()


def test_batching_config_workload(torch_cuda, tmp_path, monkeypatch):
    """The batching configuration itself: 4096 C2C problems, lengths log-uniform over 2^a 3^b 5^c in
    [64, 65536], batch U[1, 64] (257 distinct lengths, 131,586 transforms), inputs uniform[-1, 1).
    The autotuned hipGraph launch (lengths up to 4096 in the persistent multi-length launch with
    tuned radix orders) is bit-identical to the same tuned plans launched directly, and
    sampled rows of every length (the first and the last row of each length group, plus one random
    problem's rows) are within 2e-6 of the FP64 oracle for the graph and for the untuned grouped
    launches."""
    torch = torch_cuda
    monkeypatch.setenv("SIP_FFT_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    lengths, batches = S.batching_workload(4096)
    elems = [n * b for n, b in zip(lengths, batches)]
    offs = np.concatenate([[0], np.cumsum(elems)]).astype(np.int64)
    total = int(offs[-1])
    print(f"batching workload: {len(lengths)} problems, {len(set(lengths))} lengths, {total} points", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5EED0004)
    x = torch.view_as_complex(torch.rand((total, 2), generator=g, device="cuda") * 2 - 1)
    y_graph = torch.empty_like(x)
    y_direct = torch.empty_like(x)
    y_plain = torch.empty_like(x)
    base = x.data_ptr()
    ins = [base + 8 * int(o) for o in offs[:-1]]

    def outs(t):
        return [t.data_ptr() + 8 * int(o) for o in offs[:-1]]

    s = torch.cuda.current_stream().cuda_stream
    handles = []
    tuned = S.BATCH_AUTOTUNE | S.BATCH_PERSIST
    for y, flags in ((y_graph, S.BATCH_GRAPH | tuned), (y_direct, tuned), (y_plain, 0)):
        bh = S.BatchHandle()
        assert bh.make_plan(lengths, batches, ins, outs(y), FWD, flags) == 0
        bh.alloc_workspace()
        if s:
            assert bh.set_stream(s) == 0
        assert bh.exec() == 0
        handles.append(bh)
    torch.cuda.synchronize()
    d = handles[0].describe()
    assert d["groups"] + d["persist"]["groups"] == len(set(lengths)) and d["graph"] == 1
    assert d["persist"]["groups"] == len({n for n in lengths if n <= 4096})
    assert torch.equal(y_graph, y_direct), "graph launch differs from the direct launches of the same plans"
    # sampled rows: first and last row of every length group, and all rows of one random problem
    first, last = {}, {}
    for i, n in enumerate(lengths):
        first.setdefault(n, i)
        last[n] = i
    rows = []
    for n in sorted(first):
        rows.append((int(offs[first[n]]), n))
        i = last[n]
        rows.append((int(offs[i]) + (batches[i] - 1) * n, n))
    pick = int(np.random.default_rng(5).integers(0, len(lengths)))
    rows += [(int(offs[pick]) + r * lengths[pick], lengths[pick]) for r in range(batches[pick])]
    worst = 0.0
    for start, n in rows:
        xr = x[start:start + n].cpu().numpy().reshape(1, n)
        ref = oracle.c2c(xr)
        for name, y in (("graph", y_graph), ("plain grouped", y_plain)):
            rel, _ = oracle.error(y[start:start + n].cpu().numpy().reshape(1, n), ref)
            worst = max(worst, rel)
            assert rel <= TOL, f"{name} row at {start} (n={n}): rel-L2 {rel:.3e}"
    print(f"[parity] batching workload: {len(rows)} sampled rows, worst rel-L2 {worst:.3e}", flush=True)
    for bh in handles:
        bh.destroy()


# short lengths for the persistent multi-length launch (2,3,5-smooth up to 4096), mixed with lengths
# that stay group launches (1, a Bluestein length, a single-pass long leaf, a four-step length)
P_LENGTHS = [8, 12, 20, 30, 64, 100, 128, 243, 625, 1000, 1536, 2000, 3000, 4096, 4096, 5000, 20000, 1, 17, 2, 3, 64]
P_BATCHES = [5, 3, 7, 2, 9, 4, 1, 3, 2, 6, 2, 3, 1, 2, 5, 2, 1, 4, 3, 11, 6, 1]


def _smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


P_SHORT = {n for n in P_LENGTHS if 2 <= n <= 4096 and _smooth(n)}


@pytest.mark.parametrize("direction", [FWD, INV])
def test_persistent_launch_matches_alone(torch_cuda, direction):
    """The persistent launch with each length's default radix order runs the leaf kernels' passes
    (same twiddle tables, same butterflies; the compiler may contract multiply-adds differently in
    the two kernels): every problem is within 2e-6 of the FP64 oracle (5e-6 for the Bluestein length
    17, which stays a group launch) and within 5e-7 (rel-L2) of the problem alone through
    asdFftExecC2C. Graph and direct launches."""
    torch = torch_cuda
    short = P_SHORT
    for flags in (S.BATCH_PERSIST, S.BATCH_PERSIST | S.BATCH_GRAPH):
        hs, xs, ys = _problems(torch, P_LENGTHS, P_BATCHES, 700)
        bh = _run_batch(torch, P_LENGTHS, P_BATCHES, xs, ys, direction, flags)
        d = bh.describe()
        assert d["persist"]["groups"] == len(short)
        assert d["groups"] == len(set(P_LENGTHS)) - len(short)
        for g in d["persist"]["group_list"]:
            assert int(np.prod(g["radices"])) == g["n"] and g["F"] * g["T"] <= 256
        for i, (n, b) in enumerate(zip(P_LENGTHS, P_BATCHES)):
            got = ys[i].cpu().numpy()
            rel, _ = report(f"persist flags={flags} problem {i} n={n} b={b}", got,
                            oracle.c2c(hs[i], oracle.FORWARD if direction == FWD else oracle.INVERSE))
            assert rel <= (5e-6 if n == 17 else TOL)
            h = S.FftHandle()
            assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, direction, b) == 0
            h.alloc_workspace()
            h.use_current_torch_stream()
            alone = torch.empty_like(ys[i])
            assert h.exec_c2c(xs[i], alone) == 0
            assert h.synchronize() == 0
            rel_alone, _ = oracle.error(ys[i].cpu().numpy(), alone.cpu().numpy().astype(np.complex128))
            assert rel_alone <= 5e-7, f"problem {i} (n={n}): persistent vs alone rel-L2 {rel_alone:.3e}"
            h.destroy()
        bh.destroy()


def test_persistent_tuned_radix_orders(torch_cuda, tmp_path, monkeypatch):
    """Autotuned radix orders (each candidate timed alone, the fastest kept and cached): results within
    2e-6 of the oracle; the cache holds one radix line per short length; a second plan reuses it and
    its graph replays (eager work and synchronisation in between) are bit-identical."""
    torch = torch_cuda
    cache = tmp_path / "tune.jsonl"
    monkeypatch.setenv("SIP_FFT_TUNE_CACHE", str(cache))
    flags = S.BATCH_PERSIST | S.BATCH_GRAPH | S.BATCH_AUTOTUNE
    hs, xs, ys = _problems(torch, P_LENGTHS, P_BATCHES, 900)
    bh = _run_batch(torch, P_LENGTHS, P_BATCHES, xs, ys, FWD, flags)
    d = bh.describe()
    short = P_SHORT
    assert len(d["persist_tuned_us"]) == len(short) and all(u > 0 for u in d["persist_tuned_us"])
    for i, (n, b) in enumerate(zip(P_LENGTHS, P_BATCHES)):
        rel, _ = report(f"persist tuned n={n}", ys[i].cpu().numpy(), oracle.c2c(hs[i]))
        assert rel <= (5e-6 if n == 17 else TOL)
    lines = [ln for ln in cache.read_text().splitlines() if '"radices"' in ln]
    assert len(lines) == len(short)
    first = [y.clone() for y in ys]
    bh.destroy()
    bh = _run_batch(torch, P_LENGTHS, P_BATCHES, xs, ys, FWD, flags)    # decisions from the cache
    assert [g["radices"] for g in bh.describe()["persist"]["group_list"]] == \
        [g["radices"] for g in d["persist"]["group_list"]]
    for _ in range(3):
        for y in ys:
            y.zero_()
        torch.cuda.synchronize()
        assert bh.exec() == 0
        torch.cuda.synchronize()
        for a, b in zip(first, ys):
            assert torch.equal(a, b)
    bh.destroy()
