"""sip_amd: MI355X-native FFT library with the reference's asdFft* API.

The product is the C-ABI shared library ``sip_amd/libsip_fft.so`` (hand-written gfx950 HIP
kernels behind the asdFft* entry points of the reference's include/fft_api.h). This module is a
thin ctypes binding that mirrors that interface for Python callers and tests: the same entry
point names, argument meaning and status codes (see ``include/sip_fft.h``). Device memory and
streams come from PyTorch (plumbing only); every transform runs in the HIP library, and loading
fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SIP_FFT_LIB") or os.path.join(HERE, "libsip_fft.so")

# enums (reference include/fft_api.h:26-44)
ASCEND_FFT_C2C = 0x10
ASCEND_FFT_C2R = 0x11
ASCEND_FFT_R2C = 0x12
ASCEND_STFT_C2C = 0x20
ASCEND_STFT_C2R = 0x21
ASCEND_STFT_R2C = 0x22
ASCEND_FFT_C2C_SEP = 0x30
ASCEND_FFT_FORWARD = 0x10
ASCEND_FFT_INVERSE = 0x11
ASCEND_FFT_HORIZONTAL = 0x10
ASCEND_FFT_VERTICAL = 0x11

# status codes (reference core/utils/include/utils/aspb_status.h:16-27)
ACL_SUCCESS = 0
ACL_ERROR_INVALID_PARAM = 100000
ACL_ERROR_OP_INPUT_NOT_MATCH = 100021
ACL_ERROR_OP_OUTPUT_NOT_MATCH = 100022
ACL_ERROR_UNSUPPORTED_DATA_TYPE = 100026
ACL_ERROR_FORMAT_NOT_MATCH = 100027
ACL_ERROR_API_NOT_SUPPORT = 200001
ACL_ERROR_INTERNAL_ERROR = 500000

# aclDataType values used on the path
ACL_FLOAT = 0
ACL_COMPLEX64 = 16
ACL_FORMAT_ND = 2

EXPORTED_SYMBOLS = [
    "asdFftCreate", "asdFftSetStream", "asdFftMakePlan1D", "asdFftMakePlan2D", "asdFftMakePlan3D",
    "asdFftGetWorkspaceSize", "asdFftSetWorkspace", "asdFftExecC2C", "asdFftExecC2R", "asdFftExecR2C",
    "asdFftExecC2CSeparated", "asdFftIstftMakePlan", "asdFftExecIstft", "asdFftSynchronize", "asdFftGetType",
    "asdFftDestroy", "aclCreateTensor", "aclDestroyTensor", "aclGetDataType", "aclGetViewShape",
    "aclFreeViewShape", "aclGetTensorData", "sipFftPlanDescribe", "sipFftPlanDescribeDry", "sipFftVersion",
    "sipFftBatchCreate", "sipFftBatchSetStream", "sipFftBatchMakePlan", "sipFftBatchGetWorkspaceSize",
    "sipFftBatchSetWorkspace", "sipFftBatchExec", "sipFftBatchDescribe", "sipFftBatchDestroy",
    "sipFftServeCreate", "sipFftServeSubmit", "sipFftServeSubmitMany", "sipFftServeWait", "sipFftServeDrain",
    "sipFftServeStats",
    "sipFftServeBenchmark", "sipFftServeDestroy", "sipFftSetPrecision", "sipFftPlanTableDry",
    "sipFftCoeffCacheStats", "sipFftExecC2CNoAlloc",
]

PREC_FP32, PREC_F16_SPLIT, PREC_BF16_SPLIT, PREC_BF16_SPLIT3 = 0, 1, 2, 3

BATCH_GRAPH = 1
BATCH_AUTOTUNE = 2
BATCH_PERSIST = 4

_lib = None


class SipFftError(RuntimeError):
    def __init__(self, what, status):
        super().__init__(f"{what} failed with status {status}")
        self.status = status


def lib():
    """Loads libsip_fft.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sip_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
    P = ctypes.POINTER
    sig = {
        "asdFftCreate": [P(vp)],
        "asdFftSetStream": [vp, vp],
        "asdFftMakePlan1D": [vp, i64, i32, i32, i64, i32],
        "asdFftMakePlan2D": [vp, i64, i64, i32, i32, i32],
        "asdFftMakePlan3D": [vp, i64, i64, i64, i32, i32, i32],
        "asdFftGetWorkspaceSize": [vp, P(sz)],
        "asdFftSetWorkspace": [vp, vp],
        "asdFftExecC2C": [vp, vp, vp],
        "asdFftExecC2R": [vp, vp, vp],
        "asdFftExecR2C": [vp, vp, vp],
        "asdFftExecC2CSeparated": [vp, vp, vp, vp, vp],
        "asdFftIstftMakePlan": [vp, vp, i64, i64, i64, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, i64,
                                ctypes.c_bool],
        "asdFftExecIstft": [vp, vp, vp, vp],
        "asdFftSynchronize": [vp],
        "asdFftGetType": [vp, P(i32)],
        "asdFftDestroy": [vp],
        "sipFftPlanDescribe": [vp, ctypes.c_char_p, sz, P(sz)],
        "sipFftPlanDescribeDry": [i32, P(i64), i32, i32, i64, i32, ctypes.c_char_p, sz, P(sz)],
        "sipFftBatchCreate": [P(vp)],
        "sipFftBatchSetStream": [vp, vp],
        "sipFftBatchMakePlan": [vp, i32, P(i64), P(i64), P(vp), P(vp), i32, i32],
        "sipFftBatchGetWorkspaceSize": [vp, P(sz)],
        "sipFftBatchSetWorkspace": [vp, vp],
        "sipFftBatchExec": [vp],
        "sipFftBatchDescribe": [vp, ctypes.c_char_p, sz, P(sz)],
        "sipFftBatchDestroy": [vp],
        "sipFftServeCreate": [P(vp), sz, i32, i64],
        "sipFftServeSubmit": [vp, i32, i64, i64, vp, vp, P(ctypes.c_uint64)],
        "sipFftServeSubmitMany": [vp, i64, P(i32), P(i64), P(i64), P(vp), P(vp), P(ctypes.c_uint64), P(i64)],
        "sipFftServeWait": [vp, ctypes.c_uint64, P(ctypes.c_double)],
        "sipFftServeDrain": [vp],
        "sipFftServeStats": [vp, i32, ctypes.c_char_p, sz, P(sz)],
        "sipFftServeBenchmark": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, vp, sz, ctypes.c_double,
                                 i64, i64, i64, ctypes.c_double, ctypes.c_char_p, sz, P(sz)],
        "sipFftServeDestroy": [vp],
        "sipFftSetPrecision": [vp, i32],
        "sipFftPlanTableDry": [i32, P(i64), i32, i32, i64, i32, i32, i32, P(ctypes.c_float), sz, P(sz)],
        "sipFftCoeffCacheStats": [P(sz), P(sz), P(sz)],
        "sipFftExecC2CNoAlloc": [vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L.aclCreateTensor.argtypes = [P(i64), ctypes.c_uint64, ctypes.c_int, P(i64), i64, ctypes.c_int, P(i64),
                                  ctypes.c_uint64, vp]
    L.aclCreateTensor.restype = vp
    L.aclDestroyTensor.argtypes = [vp]
    L.aclDestroyTensor.restype = ctypes.c_int
    L.aclGetDataType.argtypes = [vp, P(ctypes.c_int)]
    L.aclGetDataType.restype = ctypes.c_int
    L.aclGetViewShape.argtypes = [vp, P(P(i64)), P(ctypes.c_uint64)]
    L.aclGetViewShape.restype = ctypes.c_int
    L.aclFreeViewShape.argtypes = [P(i64)]
    L.aclFreeViewShape.restype = None
    L.aclGetTensorData.argtypes = [vp]
    L.aclGetTensorData.restype = vp
    L.sipFftVersion.argtypes = []
    L.sipFftVersion.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(status, what):
    if status != ACL_SUCCESS:
        raise SipFftError(what, status)


class AclTensor:
    """aclTensor descriptor of a device buffer (address, shape, dtype); owns only the descriptor."""

    def __init__(self, data_ptr: int, shape, dtype: int):
        shape = [int(s) for s in shape]
        arr = (ctypes.c_int64 * max(1, len(shape)))(*shape)
        strides = []
        acc = 1
        for s in reversed(shape):
            strides.append(acc)
            acc *= s
        strides = list(reversed(strides))
        st = (ctypes.c_int64 * max(1, len(shape)))(*strides)
        self.ptr = lib().aclCreateTensor(arr, len(shape), dtype, st, 0, ACL_FORMAT_ND, arr, len(shape),
                                         ctypes.c_void_p(int(data_ptr)))
        if not self.ptr:
            raise RuntimeError("aclCreateTensor failed")

    @classmethod
    def from_torch(cls, t):
        import torch
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        if t.dtype == torch.complex64:
            dt = ACL_COMPLEX64
        elif t.dtype == torch.float32:
            dt = ACL_FLOAT
        else:
            raise ValueError(f"unsupported dtype {t.dtype}")
        return cls(t.data_ptr(), t.shape, dt)

    def close(self):
        if self.ptr:
            lib().aclDestroyTensor(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FftHandle:
    """One asdFftHandle: Create -> MakePlan -> workspace -> SetStream -> Exec -> Synchronize -> Destroy,
    the call order of the reference's examples (example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:121-145)."""

    _TENSOR_CACHE = 16  # descriptors kept per handle (oldest dropped first)

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().asdFftCreate(ctypes.byref(h)), "asdFftCreate")
        self.h = h
        self._ws = None
        self._tensors = {}  # (address, shape, dtype) -> AclTensor

    def _desc(self, t):
        """aclTensor descriptor of a contiguous tensor, reused across execs on the same buffer (a
        descriptor holds only address, shape and dtype, as a caller of the C API would keep it)."""
        import torch
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        key = (t.data_ptr(), tuple(t.shape), t.dtype)
        d = self._tensors.get(key)
        if d is None:
            if t.dtype not in (torch.complex64, torch.float32):
                raise ValueError(f"unsupported dtype {t.dtype}")
            if len(self._tensors) >= self._TENSOR_CACHE:
                self._tensors.pop(next(iter(self._tensors))).close()
            d = self._tensors[key] = AclTensor.from_torch(t)
        return d

    # -- planning
    def make_plan_1d(self, n, fft_type, direction, batch, dim_type=ASCEND_FFT_HORIZONTAL):
        return lib().asdFftMakePlan1D(self.h, n, fft_type, direction, batch, dim_type)

    def make_plan_2d(self, x, y, fft_type, direction, batch):
        return lib().asdFftMakePlan2D(self.h, x, y, fft_type, direction, batch)

    def make_plan_3d(self, x, y, z, fft_type, direction, batch):
        return lib().asdFftMakePlan3D(self.h, x, y, z, fft_type, direction, batch)

    def set_precision(self, mode):
        return lib().sipFftSetPrecision(self.h, mode)

    def workspace_size(self) -> int:
        n = ctypes.c_size_t()
        _check(lib().asdFftGetWorkspaceSize(self.h, ctypes.byref(n)), "asdFftGetWorkspaceSize")
        return n.value

    def alloc_workspace(self, device="cuda"):
        """Allocates the workspace with torch (device memory) and binds it to the handle."""
        import torch
        n = self.workspace_size()
        if n > 0:
            self._ws = torch.empty(n, dtype=torch.uint8, device=device)
            _check(lib().asdFftSetWorkspace(self.h, ctypes.c_void_p(self._ws.data_ptr())), "asdFftSetWorkspace")
        return n

    def set_stream(self, stream_ptr: int):
        return lib().asdFftSetStream(self.h, ctypes.c_void_p(int(stream_ptr)))

    def use_current_torch_stream(self):
        import torch
        s = torch.cuda.current_stream().cuda_stream
        if s:
            _check(self.set_stream(s), "asdFftSetStream")

    # -- execution on torch tensors (no copies; descriptors only)
    def _exec(self, fn, x, y):
        return fn(self.h, ctypes.c_void_p(self._desc(x).ptr), ctypes.c_void_p(self._desc(y).ptr))

    def exec_c2c(self, x, y):
        return self._exec(lib().asdFftExecC2C, x, y)

    def exec_c2c_noalloc(self, x, y):
        """sipFftExecC2CNoAlloc: C2C with the workspace taken from the library's per-device scratch."""
        return self._exec(lib().sipFftExecC2CNoAlloc, x, y)

    def exec_r2c(self, x, y):
        return self._exec(lib().asdFftExecR2C, x, y)

    def exec_c2r(self, x, y):
        return self._exec(lib().asdFftExecC2R, x, y)

    def istft_make_plan(self, x, n_fft, hop, win_length, center=True, normalized=False, onesided=False, length=0,
                        return_complex=True):
        """asdFftIstftMakePlan for the spectra tensor x (channel, n_fft, frames), complex64."""
        t = AclTensor.from_torch(x)
        try:
            return lib().asdFftIstftMakePlan(self.h, ctypes.c_void_p(t.ptr), n_fft, hop, win_length, center,
                                             normalized, onesided, length, return_complex)
        finally:
            t.close()

    def exec_istft(self, x, window, y):
        """asdFftExecIstft: spectra x, float window (n_fft), complex output (channel, out_len)."""
        ts = [AclTensor.from_torch(t) for t in (x, window, y)]
        try:
            return lib().asdFftExecIstft(self.h, *[ctypes.c_void_p(t.ptr) for t in ts])
        finally:
            for t in ts:
                t.close()

    def exec_c2c_separated(self, x_re, x_im, y_re, y_im):
        """asdFftExecC2CSeparated: float32 real / imaginary planes of one shape each."""
        ts = [AclTensor.from_torch(t) for t in (x_re, x_im, y_re, y_im)]
        try:
            return lib().asdFftExecC2CSeparated(self.h, *[ctypes.c_void_p(t.ptr) for t in ts])
        finally:
            for t in ts:
                t.close()

    def synchronize(self):
        return lib().asdFftSynchronize(self.h)

    def get_type(self):
        t = ctypes.c_int32()
        st = lib().asdFftGetType(self.h, ctypes.byref(t))
        return st, t.value

    def describe(self) -> dict:
        return json.loads(_describe(lambda buf, cap, need: lib().sipFftPlanDescribe(self.h, buf, cap, need)))

    def destroy(self):
        if self.h is not None:
            st = lib().asdFftDestroy(self.h)
            self.h = None
            self._ws = None
            for d in self._tensors.values():
                d.close()
            self._tensors = {}
            return st
        return ACL_SUCCESS

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _describe(call) -> str:
    need = ctypes.c_size_t()
    call(None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    st = call(buf, need.value, ctypes.byref(need))
    _check(st, "sipFftPlanDescribe")
    return buf.value.decode()


def _run_report(call, cap=1 << 20) -> str:
    """One call into an entry point that runs work and reports JSON (it must not run twice)."""
    need = ctypes.c_size_t()
    buf = ctypes.create_string_buffer(cap)
    st = call(buf, cap, ctypes.byref(need))
    _check(st, "report")
    return buf.value.decode()


def describe_dry(dims, fft_type, direction, batch, dim_type=ASCEND_FFT_HORIZONTAL) -> dict:
    """Host-only plan layout (no GPU): stages, kernels, workspace, and the reference's step list."""
    dims = [int(d) for d in dims]
    arr = (ctypes.c_int64 * len(dims))(*dims)
    return json.loads(_describe(lambda buf, cap, need: lib().sipFftPlanDescribeDry(
        len(dims), arr, fft_type, direction, batch, dim_type, buf, cap, need)))


def plan_table_dry(dims, fft_type, direction, batch, stage, which, dim_type=ASCEND_FFT_HORIZONTAL):
    """Host-generated coefficient table `which` (0 pass twiddles, 1 real roots, 2 / 3 four-step low /
    high, 4 auxiliary, 5 fused-2D column twiddles) of stage `stage` of a host-only plan, as complex64."""
    import numpy as np
    dims = [int(d) for d in dims]
    arr = (ctypes.c_int64 * len(dims))(*dims)
    cnt = ctypes.c_size_t()
    st = lib().sipFftPlanTableDry(len(dims), arr, fft_type, direction, batch, dim_type, stage, which, None, 0,
                                  ctypes.byref(cnt))
    if cnt.value == 0:
        _check(st, "sipFftPlanTableDry")
    out = np.empty(2 * cnt.value, dtype=np.float32)
    _check(lib().sipFftPlanTableDry(len(dims), arr, fft_type, direction, batch, dim_type, stage, which,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size, ctypes.byref(cnt)),
           "sipFftPlanTableDry")
    return out.view(np.complex64)


def coeff_cache_stats():
    """(map entries, live tables, live bytes) of the device-resident coefficient cache."""
    e, n, b = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().sipFftCoeffCacheStats(ctypes.byref(e), ctypes.byref(n), ctypes.byref(b)), "sipFftCoeffCacheStats")
    return e.value, n.value, b.value


class BatchHandle:
    """Batching form (sipFftBatch*): uneven C2C problems bound to device buffers, one launch per exec."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().sipFftBatchCreate(ctypes.byref(h)), "sipFftBatchCreate")
        self.h = h
        self._ws = None

    def make_plan(self, lengths, batches, in_ptrs, out_ptrs, direction=ASCEND_FFT_FORWARD, flags=BATCH_GRAPH):
        n = len(lengths)
        L = (ctypes.c_int64 * n)(*[int(v) for v in lengths])
        B = (ctypes.c_int64 * n)(*[int(v) for v in batches])
        I = (ctypes.c_void_p * n)(*[int(v) for v in in_ptrs])
        O = (ctypes.c_void_p * n)(*[int(v) for v in out_ptrs])
        return lib().sipFftBatchMakePlan(self.h, n, L, B, I, O, direction, flags)

    def workspace_size(self):
        n = ctypes.c_size_t()
        _check(lib().sipFftBatchGetWorkspaceSize(self.h, ctypes.byref(n)), "sipFftBatchGetWorkspaceSize")
        return n.value

    def alloc_workspace(self, device="cuda"):
        import torch
        n = self.workspace_size()
        if n > 0:
            self._ws = torch.empty(n, dtype=torch.uint8, device=device)
            _check(lib().sipFftBatchSetWorkspace(self.h, ctypes.c_void_p(self._ws.data_ptr())),
                   "sipFftBatchSetWorkspace")
        return n

    def set_stream(self, stream_ptr):
        return lib().sipFftBatchSetStream(self.h, ctypes.c_void_p(int(stream_ptr)))

    def exec(self):
        return lib().sipFftBatchExec(self.h)

    def describe(self) -> dict:
        return json.loads(_describe(lambda buf, cap, need: lib().sipFftBatchDescribe(self.h, buf, cap, need)))

    def destroy(self):
        if self.h is not None:
            st = lib().sipFftBatchDestroy(self.h)
            self.h = None
            self._ws = None
            return st
        return ACL_SUCCESS

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def smooth_lengths(lo, hi):
    """All 2^a 3^b 5^c in [lo, hi], ascending."""
    out = []
    a = 1
    while a <= hi:
        b = a
        while b <= hi:
            c = b
            while c <= hi:
                if c >= lo:
                    out.append(c)
                c *= 5
            b *= 3
        a *= 2
    return sorted(out)


def batching_workload(count=4096, seed=0x5EED0004, lo=64, hi=65536, max_batch=64):
    """The batching-form problem set: `count` C2C problems, length log-uniform over the 2,3,5-smooth
    numbers in [lo, hi] (x = exp(U[ln lo, ln hi]), rounded up to the next smooth length), batch
    uniform in [1, max_batch]; seeded (numpy PCG64)."""
    import math
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    sm = smooth_lengths(lo, hi)
    lengths, batches = [], []
    for _ in range(count):
        x = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        n = next(v for v in sm if v >= x) if x <= hi else hi
        lengths.append(int(n))
        batches.append(int(rng.integers(1, max_batch + 1)))
    return lengths, batches


class Server:
    """Serving form (sipFftServe*): continuous-batching FFT request loop on one GPU."""

    def __init__(self, budget_bytes=64 << 30, nstreams=4, max_coalesce_rows=8192):
        h = ctypes.c_void_p()
        _check(lib().sipFftServeCreate(ctypes.byref(h), int(budget_bytes), nstreams, max_coalesce_rows),
               "sipFftServeCreate")
        self.h = h

    def submit(self, fft_type, n, batch, in_ptr, out_ptr=None):
        t = ctypes.c_uint64()
        _check(lib().sipFftServeSubmit(self.h, fft_type, n, batch, ctypes.c_void_p(int(in_ptr)),
                                       ctypes.c_void_p(int(out_ptr) if out_ptr else None), ctypes.byref(t)),
               "sipFftServeSubmit")
        return t.value

    def submit_many(self, requests):
        """requests: (fft_type, n, batch, in_ptr, out_ptr or None) tuples, queued together (all or
        none); returns their tickets."""
        k = len(requests)
        types = (ctypes.c_int32 * k)(*[int(r[0]) for r in requests])
        ns = (ctypes.c_int64 * k)(*[int(r[1]) for r in requests])
        bs = (ctypes.c_int64 * k)(*[int(r[2]) for r in requests])
        ins = (ctypes.c_void_p * k)(*[int(r[3]) for r in requests])
        outs = (ctypes.c_void_p * k)(*[int(r[4]) if r[4] else None for r in requests])
        tks = (ctypes.c_uint64 * k)()
        bad = ctypes.c_int64(-1)
        st = lib().sipFftServeSubmitMany(self.h, k, types, ns, bs, ins, outs, tks, ctypes.byref(bad))
        if st != ACL_SUCCESS:
            raise SipFftError(f"sipFftServeSubmitMany (request {bad.value})", st)
        return list(tks)

    def wait(self, ticket):
        lat = ctypes.c_double()
        st = lib().sipFftServeWait(self.h, ticket, ctypes.byref(lat))
        return st, lat.value

    def drain(self):
        return lib().sipFftServeDrain(self.h)

    def stats(self, reset=False) -> dict:
        return json.loads(_run_report(lambda buf, cap, need: lib().sipFftServeStats(self.h, int(reset), buf, cap, need)))

    def benchmark(self, rate, seconds, pool_ptr, pool_bytes, seed=1, c2c_frac=0.7, n_lo=256, n_hi=32768, b_max=512,
                  drain_seconds=30.0) -> dict:
        return json.loads(_run_report(lambda buf, cap, need: lib().sipFftServeBenchmark(
            self.h, rate, seconds, seed, ctypes.c_void_p(int(pool_ptr)), int(pool_bytes), c2c_frac, n_lo, n_hi,
            b_max, drain_seconds, buf, cap, need)))

    def destroy(self):
        if self.h is not None:
            st = lib().sipFftServeDestroy(self.h)
            self.h = None
            return st
        return ACL_SUCCESS

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
