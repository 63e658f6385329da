"""Benchmark of the FFT hot path on one MI355X.

Workload (the metric's headline configuration): batched 1D C2C complex64 forward FFT, N = 4096,
batch = 16384, inputs uniform[-1,1) already resident in HBM, one asdFftExecC2C per step (one kernel
launch). value = 5 N log2 N * batch * steps / wall time (GFLOP/s), timed over `--steps` steps
bracketed by device synchronisation after `--warmup` untimed steps.

roofline: algorithmic bytes per launch (one read + one write of 16384 x 4096 complex64 = 1 GiB)
divided by the kernel's average duration, measured with HIP events recorded on the plan's stream
around the timed steps; peak = 8 TB/s HBM. traffic (HBM bytes per launch from rocprofv3 PMC
FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel's reading) is read from
profiles/traffic_c2c_4096x16384.json when present.

cpu_baseline: the oracle's FP32 CPU FFT (same staged Stockham algorithm) on the host cores (every core
the process may use, at most 16), median of runs over the full 16384 x 4096 batch, and the FP64
oracle FFT on 4096 of the rows; the leg also checks a sample of the GPU's output rows against the
FP64 oracle, and times scipy.fft (pocketfft) on the same threads as a comparator when installed.
Secondary configurations (R2C/C2R 8192 x 4096, C2C 1536 / 2000 x 8192, C2C 1024 x 256,
2D 1024^2 x 8; and 1 GiB arrays beyond the last-level cache: four-step 65536 x 2048, 2D 1024^2 x 128,
3D 128^3 x 64) are timed the same way and reported under "secondary"; the precision, batching and
serving forms have legs of their own.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-batching", action="store_true", help="skip the uneven-batch (batching form) leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-form leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-form (MFMA) leg")
    ap.add_argument("--serving-seconds", type=float, default=60.0,
                    help="arrival window per rate (the serving configuration's 60 s by default)")
    return ap.parse_args()


def timed_exec(torch, h, kind, x, y, steps, warmup, stream):
    fn = {"c2c": h.exec_c2c, "r2c": h.exec_r2c, "c2r": h.exec_c2r}[kind]
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            assert fn(x, y) == 0
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            assert fn(x, y) == 0
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    dev_ms = ev0.elapsed_time(ev1)
    return wall, dev_ms / steps


def make_handle(S, torch, dims, fft_type, batch, stream, precision=0):
    h = S.FftHandle()
    if precision:
        assert h.set_precision(precision) == 0
    if len(dims) == 1:
        st = h.make_plan_1d(dims[0], fft_type, S.ASCEND_FFT_FORWARD, batch)
    elif len(dims) == 2:
        st = h.make_plan_2d(dims[0], dims[1], fft_type, S.ASCEND_FFT_FORWARD, batch)
    else:
        st = h.make_plan_3d(dims[0], dims[1], dims[2], fft_type, S.ASCEND_FFT_FORWARD, batch)
    assert st == 0, st
    h.alloc_workspace()
    assert h.set_stream(stream.cuda_stream) == 0
    return h


def bench_batching(torch, S, stream, reps=5):
    """Batching form: 4096 C2C problems, length log-uniform over 2^a 3^b 5^c in [64, 65536], batch
    U[1, 64] (sip_amd.batching_workload). The form: one hipGraph launch of the length-grouped plan
    with every length up to 4096 in one persistent multi-length launch, radix orders and tile shapes
    autotuned. Compared with (a) one plain asdFftExecC2C per problem, (b) per-size plain launches
    (one gathered-row launch per distinct length, default tiles, no graph: the headline ratio) and
    (c) the graph and tuner without the persistent launch."""
    import tempfile
    lengths, batches = S.batching_workload(4096)
    elems = [n * b for n, b in zip(lengths, batches)]
    total = sum(elems)
    flops = sum(5.0 * n * math.log2(n) * b for n, b in zip(lengths, batches))
    nbytes = 16.0 * total
    offs = [0]
    for e in elems:
        offs.append(offs[-1] + e)
    with torch.cuda.stream(stream):
        xin = torch.view_as_complex(torch.rand((total, 2), device="cuda") * 2 - 1).contiguous()
        yout = torch.empty_like(xin)
    torch.cuda.synchronize()
    base_in, base_out = xin.data_ptr(), yout.data_ptr()
    ins = [base_in + 8 * o for o in offs[:-1]]
    outs = [base_out + 8 * o for o in offs[:-1]]

    def run(fn):
        with torch.cuda.stream(stream):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    res = {"problems": len(lengths), "distinct_lengths": len(set(lengths)), "transforms": sum(batches),
           "algorithmic_bytes": int(nbytes), "flops_5NlogN": flops}
    # (a) one plain plan + exec per problem
    hs = []
    for n, b in zip(lengths, batches):
        h = S.FftHandle()
        assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, S.ASCEND_FFT_FORWARD, b) == 0
        h.alloc_workspace()
        assert h.set_stream(stream.cuda_stream) == 0
        hs.append(h)
    tens_in = [xin[offs[i]:offs[i + 1]].view(batches[i], lengths[i]) for i in range(len(lengths))]
    tens_out = [yout[offs[i]:offs[i + 1]].view(batches[i], lengths[i]) for i in range(len(lengths))]

    def per_problem():
        for h, a, b in zip(hs, tens_in, tens_out):
            assert h.exec_c2c(a, b) == 0
    ms_a = run(per_problem)
    for h in hs:
        h.destroy()
    # (b) grouped, direct launches, default tiles
    bh = S.BatchHandle()
    assert bh.make_plan(lengths, batches, ins, outs, S.ASCEND_FFT_FORWARD, 0) == 0
    bh.alloc_workspace()
    assert bh.set_stream(stream.cuda_stream) == 0
    def exec_b():
        assert bh.exec() == 0
    ms_b = run(exec_b)
    bh.destroy()
    # (c) grouped + autotuned + one hipGraph launch, with and without the persistent launch
    cache = os.path.join(tempfile.gettempdir(), "sip_fft_bench_tune.jsonl")
    os.environ["SIP_FFT_TUNE_CACHE"] = cache
    timed = {}
    for name, flags in (("graph_autotuned_no_persist", S.BATCH_GRAPH | S.BATCH_AUTOTUNE),
                        ("graph_autotuned_persistent", S.BATCH_GRAPH | S.BATCH_AUTOTUNE | S.BATCH_PERSIST)):
        t0 = time.perf_counter()
        bg = S.BatchHandle()
        assert bg.make_plan(lengths, batches, ins, outs, S.ASCEND_FFT_FORWARD, flags) == 0
        tune_s = time.perf_counter() - t0
        bg.alloc_workspace()
        assert bg.set_stream(stream.cuda_stream) == 0

        def exec_c():
            assert bg.exec() == 0
        timed[name] = (run(exec_c), tune_s, bg.describe())
        bg.destroy()
    ms_c, tune_s, d = timed["graph_autotuned_persistent"]
    for name, ms in (("per_problem_plain", ms_a), ("per_size_plain", ms_b),
                     ("graph_autotuned_no_persist", timed["graph_autotuned_no_persist"][0]),
                     ("graph_autotuned_persistent", ms_c)):
        res[name] = {"ms": round(ms, 4), "GFLOP/s": round(flops / (ms * 1e-3) / 1e9, 1),
                     "hbm_GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    res["speedup_vs_per_size_plain"] = round(ms_b / ms_c, 3)
    res["speedup_vs_per_problem_plain"] = round(ms_a / ms_c, 3)
    res["group_launches"] = d["groups"]
    res["persistent"] = {k: d["persist"][k] for k in ("groups", "tiles", "grid", "lds")} if "persist" in d else None
    res["chains"] = d["chains"]
    # the first tuned plan tunes every group and caches the decisions; the second reuses them and
    # tunes the radix orders of the persistent launch's lengths
    res["plan_and_tune_s"] = round(timed["graph_autotuned_no_persist"][1] + tune_s, 2)
    del xin, yout
    torch.cuda.empty_cache()
    return res


def bench_precision(torch, S, stream, steps):
    """Precision form next to the plain form: the same C2C plans with the radix-16 stages on the
    matrix cores (FP16 hi+lo, BF16 hi+lo, BF16 hi+mid+lo), GFLOP/s (5 N log2 N) and device time per
    exec. Inputs: uniform[-1, 1) for the 1D headline shape (rel-L2 / max-abs vs the FP64 oracle on 64
    sampled rows); the precision configuration's Gaussian N(0,1) + 3 tones of amplitude 1e3 at random
    bins for the 2D 1024^2 x 8 and 3D 128^3 x 4 cases (rel-L2 / max-abs vs the FP64 oracle over the
    whole array), per dtype."""
    import numpy as np
    import oracle
    modes = [("fp32", 0), ("f16x2", S.PREC_F16_SPLIT), ("bf16x2", S.PREC_BF16_SPLIT),
             ("bf16x3", S.PREC_BF16_SPLIT3)]
    cases = [("c2c_4096x16384", [4096], 16384, None), ("c2c_2d_1024x1024x8", [1024, 1024], 8, 0x5EED0010),
             ("c2c_3d_128x128x128x4", [128, 128, 128], 4, 0x5EED0011)]
    res = {}
    for name, dims, b, tone_seed in cases:
        nn = int(math.prod(dims))
        if tone_seed is None:
            with torch.cuda.stream(stream):
                g = torch.Generator(device="cuda")
                g.manual_seed(0x5EED0005)
                xi = torch.view_as_complex(torch.rand([b] + dims + [2], generator=g, device="cuda") * 2 - 1).contiguous()
            torch.cuda.synchronize()  # xi was written on `stream`; the host copy below runs on another
            rows = np.arange(0, b, b // 64)
            xs = xi.cpu().numpy()[rows]
            ref = oracle.c2c(xs)
        else:
            xh = oracle.tones_complex([b] + dims, seed=tone_seed)
            rows = None
            ref = oracle.c2c_nd(xh, len(dims))
            xi = torch.from_numpy(xh).to("cuda")
        with torch.cuda.stream(stream):
            yo = torch.empty_like(xi)
        torch.cuda.synchronize()
        fl = 5.0 * nn * math.log2(nn) * b
        entry = {"input": "uniform[-1,1), 64 sampled rows checked" if rows is not None
                 else "N(0,1) + 3 tones of 1e3 at random bins, whole array checked"}
        for mname, mode in modes:
            hh = make_handle(S, torch, dims, S.ASCEND_FFT_C2C, b, stream, mode)
            st = max(10, steps // 2)
            w, kms = timed_exec(torch, hh, "c2c", xi, yo, st, 3, stream)
            e = {"GFLOP/s": round(fl * st / w / 1e9, 1), "device_ms": round(kms, 5),
                 "device_GFLOP/s": round(fl / (kms * 1e-3) / 1e9, 1)}
            got = yo.cpu().numpy()
            rel, mx = oracle.error(got[rows] if rows is not None else got, ref)
            e["rel_l2"] = float(f"{rel:.3e}")
            e["max_abs"] = float(f"{mx:.3e}")
            entry[mname] = e
            hh.destroy()
        res[name] = entry
        del xi, yo
    return res


def bench_serving(torch, S, seconds, rates=(2000.0, 10000.0, 50000.0)):
    """Serving form: open-loop Poisson streams at the configured rates (70 % C2C / 30 % R2C, length
    log-uniform over 2^a 3^b 5^c in [256, 32768], batch U[1, 512], inputs uniform[-1,1) slices of a
    device pool) into a server with a 64 GB admission budget, 4 streams and coalescing. Each rate
    submits for `seconds`, then gets the same time to drain; requests still unadmitted are dropped
    and counted. A sustainable rate (about half the measured capacity) is reported as well."""
    srv = S.Server(budget_bytes=64 << 30, nstreams=4, max_coalesce_rows=8192)
    pool = torch.rand(256 << 20, device="cuda") * 2 - 1          # 1 GiB pool
    torch.cuda.synchronize()
    res = {"config": {"c2c_frac": 0.7, "n": [256, 32768], "batch": [1, 512], "budget_GB": 64, "streams": 4,
                      "window_s": seconds}}
    # Cold start: the first request of each (type, n) builds and uploads that shape's coefficient
    # tables (the server then keeps them resident). A one-second stream at a moderate rate is served
    # first and reported on its own, so the rates below measure the steady state.
    print("[bench] serving: cold start (5000 req/s, 1 s)", file=sys.stderr, flush=True)
    r = srv.benchmark(rate=5000.0, seconds=1.0, pool_ptr=pool.data_ptr(), pool_bytes=pool.numel() * 4,
                      seed=7, drain_seconds=5.0)
    res["cold_start_5000"] = {"completed": r["completed"], "p50_ms": round(r["server"]["p50_ms"], 2),
                              "p99_ms": round(r["server"]["p99_ms"], 2),
                              "note": "first 1 s stream of a fresh server: includes building every shape's tables"}
    cap = None
    for i, rate in enumerate(rates):
        print(f"[bench] serving: {rate:.0f} req/s for {seconds:.0f} s", file=sys.stderr, flush=True)
        r = srv.benchmark(rate=rate, seconds=seconds, pool_ptr=pool.data_ptr(), pool_bytes=pool.numel() * 4,
                          seed=100 + i, drain_seconds=seconds)
        sv = r["server"]
        busy = max(r["busy_s"], 1e-9)
        res[f"rate_{int(rate)}"] = {"offered_req_s": rate, "submitted": r["submitted"], "completed": r["completed"],
                                     "dropped_unadmitted": r["dropped_unadmitted"],
                                     "achieved_req_s": round(r["achieved_req_per_s"], 1),
                                     "p50_ms": round(sv["p50_ms"], 2), "p99_ms": round(sv["p99_ms"], 2),
                                     "GB/s": round(sv["bytes_done"] / busy / 1e9, 1),
                                     "GFLOP/s": round(sv["flops_done"] / busy / 1e9, 1),
                                     "launches": sv["launches"], "coalesced_requests": sv["coalesced_requests"],
                                     "arena_peak_GB": round(sv["arena_peak"] / 1e9, 2)}
        cap = max(cap or 0.0, r["achieved_req_per_s"])
    sustain = max(1.0, 0.5 * (cap or 2.0))
    sus_s = min(seconds, 10.0)
    print(f"[bench] serving: sustainable {sustain:.0f} req/s for {sus_s:.0f} s", file=sys.stderr, flush=True)
    r = srv.benchmark(rate=sustain, seconds=sus_s, pool_ptr=pool.data_ptr(), pool_bytes=pool.numel() * 4,
                      seed=99, drain_seconds=2 * sus_s)
    sv = r["server"]
    res["sustainable"] = {"offered_req_s": round(sustain, 1), "completed": r["completed"],
                          "dropped_unadmitted": r["dropped_unadmitted"], "p50_ms": round(sv["p50_ms"], 2),
                          "p99_ms": round(sv["p99_ms"], 2), "achieved_req_s": round(r["achieved_req_per_s"], 1)}
    srv.destroy()
    del pool
    torch.cuda.empty_cache()
    return res


def settle(torch, leg):
    """Surfaces a pending device error right after the leg that left it (a small launch checks it)."""
    torch.cuda.synchronize()
    try:
        torch.zeros(1, device="cuda").add_(1)
        torch.cuda.synchronize()
    except Exception as e:  # noqa: BLE001
        raise RuntimeError(f"device error pending after the {leg} leg: {e}") from e


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (the batched 1D path "
                                   "would shard by splitting the batch across GPUs, one process each)"}))
        sys.exit(2)
    import torch
    import sip_amd as S

    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    n, batch = 4096, 16384
    flops = 5.0 * n * math.log2(n) * batch
    alg_bytes = 16.0 * n * batch
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5EED0002)
    with torch.cuda.stream(stream):
        x = (torch.rand((batch, n, 2), generator=g, device="cuda", dtype=torch.float32) * 2 - 1)
        x = torch.view_as_complex(x).contiguous()
        y = torch.empty_like(x)
    torch.cuda.synchronize()
    h = make_handle(S, torch, [n], S.ASCEND_FFT_C2C, batch, stream)
    wall, kern_ms = timed_exec(torch, h, "c2c", x, y, args.steps, args.warmup, stream)
    value = flops * args.steps / wall / 1e9
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2c_4096x16384.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    # device-to-device copy of the same 512 MiB (read + write = the FFT's algorithmic bytes): the
    # practical HBM ceiling next to the 8 TB/s peak
    with torch.cuda.stream(stream):
        z = torch.empty_like(x)
        for _ in range(3):
            z.copy_(x)
        c0 = torch.cuda.Event(enable_timing=True)
        c1 = torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        for _ in range(10):
            z.copy_(x)
        c1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = alg_bytes / (c0.elapsed_time(c1) / 10 * 1e-3) / 1e9
    del z

    out = {
        "metric": "FFT GFLOP/s (5N log2N) + %HBM roofline at rel-L2<=2e-6; serving p50/p99 ms",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (uniform[-1,1) complex64, seeded torch generator)",
        "config": {"workload": "plain batched 1D C2C forward, complex64, N=4096, batch=16384",
                   "n": n, "batch": batch, "direction": "forward"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel_ms": round(kern_ms, 5), "algorithmic_bytes": int(alg_bytes),
                     "copy_GB/s": round(copy_gbs, 1)},
    }

    if not args.no_secondary:
        print("[bench] secondary", file=sys.stderr, flush=True)
        sec = {}
        cases = [("c2c_1024x256", [1024], S.ASCEND_FFT_C2C, 256),
                 ("c2c_1536x8192", [1536], S.ASCEND_FFT_C2C, 8192),
                 ("c2c_2000x8192", [2000], S.ASCEND_FFT_C2C, 8192),
                 ("r2c_8192x4096", [8192], S.ASCEND_FFT_R2C, 4096),
                 ("c2r_8192x4096", [8192], S.ASCEND_FFT_C2R, 4096),
                 ("c2c_2d_1024x1024x8", [1024, 1024], S.ASCEND_FFT_C2C, 8),
                 ("c2c_65536x2048", [65536], S.ASCEND_FFT_C2C, 2048),
                 ("c2c_2d_1024x1024x128", [1024, 1024], S.ASCEND_FFT_C2C, 128),
                 ("c2c_3d_128x128x128x64", [128, 128, 128], S.ASCEND_FFT_C2C, 64)]
        for name, dims, ft, b in cases:
            nn = int(math.prod(dims))
            with torch.cuda.stream(stream):
                if ft == S.ASCEND_FFT_R2C:
                    xi = torch.rand((b, nn), device="cuda") * 2 - 1
                    yo = torch.empty((b, nn // 2 + 1), dtype=torch.complex64, device="cuda")
                    kind, fl, by = "r2c", 2.5 * nn * math.log2(nn) * b, 4.0 * nn * b + 8.0 * (nn // 2 + 1) * b
                elif ft == S.ASCEND_FFT_C2R:
                    xi = torch.view_as_complex(torch.rand((b, nn // 2 + 1, 2), device="cuda")).contiguous()
                    yo = torch.empty((b, nn), dtype=torch.float32, device="cuda")
                    kind, fl, by = "c2r", 2.5 * nn * math.log2(nn) * b, 4.0 * nn * b + 8.0 * (nn // 2 + 1) * b
                else:
                    shape = [b] + dims
                    xi = torch.view_as_complex(torch.rand(shape + [2], device="cuda") * 2 - 1).contiguous()
                    yo = torch.empty_like(xi)
                    kind, fl, by = "c2c", 5.0 * nn * math.log2(nn) * b, 16.0 * nn * b
            torch.cuda.synchronize()
            hh = make_handle(S, torch, dims, ft, b, stream)
            st = max(10, args.steps // 2)
            w, kms = timed_exec(torch, hh, kind, xi, yo, st, 3, stream)
            sec[name] = {"GFLOP/s": round(fl * st / w / 1e9, 1), "ms_per_exec": round(w / st * 1e3, 4),
                         "device_ms": round(kms, 5), "hbm_GB/s": round(by / (kms * 1e-3) / 1e9, 1),
                         "hbm_frac": round(by / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                         "flop_convention": "2.5N log2N" if kind != "c2c" else "5N log2N"}
            hh.destroy()
            del xi, yo
        out["secondary"] = sec
        settle(torch, "secondary")

    if not args.no_precision:
        print("[bench] precision", file=sys.stderr, flush=True)
        out["precision"] = bench_precision(torch, S, stream, args.steps)
        settle(torch, "precision")

    if not args.no_batching:
        print("[bench] batching", file=sys.stderr, flush=True)
        out["batching"] = bench_batching(torch, S, stream)
        settle(torch, "batching")

    if not args.no_serving:
        print("[bench] serving", file=sys.stderr, flush=True)
        out["serving"] = bench_serving(torch, S, args.serving_seconds)

    if not args.no_cpu:
        import numpy as np
        import oracle
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        visible = len(os.sched_getaffinity(0))
        # worker threads: every core this process may use, capped at 16 (the host share of one GPU
        # on an MI355X node; OMP_NUM_THREADS on the box says the same)
        threads = min(visible, 16)
        xs = x.cpu().numpy()
        ys = y.cpu().numpy()
        oracle.c2c_f32(xs[:256], threads=threads)  # warm-up
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.c2c_f32(xs, threads=threads)
            times.append(time.perf_counter() - t0)
        tmed = sorted(times)[len(times) // 2]
        # FP64 oracle FFT (OpenMP over the batch) on a bounded sample: 4096 of the 16384 rows
        sample = xs[:4096]
        oracle.c2c(sample[:64])
        t64 = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.c2c(sample)
            t64.append(time.perf_counter() - t0)
        t64m = sorted(t64)[1]
        fl64 = 5.0 * n * math.log2(n) * sample.shape[0]
        rows = np.arange(0, batch, batch // 64)
        rel, mx = oracle.error(ys[rows], oracle.c2c(xs[rows]))
        out["cpu_baseline"] = {"value": round(flops / tmed / 1e9, 2), "unit": "GFLOP/s", "cores": threads,
                               "cores_visible": visible, "omp_threads_fp64": oracle.max_threads(),
                               "kind": "port",
                               "sample": "FP32: full C2C N=4096 x batch 16384, staged Stockham (oracle/fft_oracle.c), "
                                         "OpenMP over the batch on `cores` threads, median of 5 runs; FP64: the "
                                         "oracle's FFT on the first 4096 rows, median of 3",
                               "seconds_per_batch": round(tmed, 3),
                               "fp64_value": round(fl64 / t64m / 1e9, 2), "fp64_seconds_4096_rows": round(t64m, 3),
                               "gpu_rel_l2_vs_fp64_sample": rel, "gpu_max_abs_vs_fp64_sample": mx,
                               "gpu_check_rows": int(rows.size)}
        # extra comparator: pocketfft (scipy.fft, single precision for complex64) on the same threads
        try:
            import scipy.fft as sfft
        except ImportError:
            sfft = None
        if sfft is not None:
            sfft.fft(xs[:256], axis=-1, workers=threads)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                sfft.fft(xs, axis=-1, workers=threads)
                ts.append(time.perf_counter() - t0)
            tsm = sorted(ts)[1]
            out["cpu_baseline"].update({"scipy_fft_value": round(flops / tsm / 1e9, 2),
                                        "scipy_fft_seconds_per_batch": round(tsm, 3),
                                        "scipy_fft_note": "comparator only: scipy.fft (pocketfft) complex64, "
                                                          "workers = `cores`, full batch, median of 3"})
    h.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
