"""Shared helpers for the GPU parity tests: run a transform through the C-ABI library on torch
device memory, and compare with the FP64 oracle (rel-L2 and max-abs)."""
from __future__ import annotations

import numpy as np

import oracle
import sip_amd as S

FWD, INV = S.ASCEND_FFT_FORWARD, S.ASCEND_FFT_INVERSE


def sign_of(direction):
    return oracle.FORWARD if direction == FWD else oracle.INVERSE


def _torch():
    import torch
    return torch


def to_dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def make(dims, fft_type, direction, batch, dim_type=S.ASCEND_FFT_HORIZONTAL, precision=0):
    h = S.FftHandle()
    if precision:
        assert h.set_precision(precision) == S.ACL_SUCCESS
    if len(dims) == 1:
        st = h.make_plan_1d(dims[0], fft_type, direction, batch, dim_type)
    elif len(dims) == 2:
        st = h.make_plan_2d(dims[0], dims[1], fft_type, direction, batch)
    else:
        st = h.make_plan_3d(dims[0], dims[1], dims[2], fft_type, direction, batch)
    assert st == S.ACL_SUCCESS, f"make plan {dims} type {fft_type:#x} -> {st}"
    h.alloc_workspace()
    h.use_current_torch_stream()
    return h


def run(h, kind, x_host, out_shape, out_dtype):
    torch = _torch()
    xd = to_dev(x_host)
    yd = torch.empty(out_shape, dtype=out_dtype, device="cuda")
    fn = {"c2c": h.exec_c2c, "r2c": h.exec_r2c, "c2r": h.exec_c2r}[kind]
    st = fn(xd, yd)
    assert st == S.ACL_SUCCESS, f"exec {kind} -> {st}"
    assert h.synchronize() == S.ACL_SUCCESS
    # the input is not modified
    assert torch.equal(xd.cpu(), torch.from_numpy(np.ascontiguousarray(x_host))), "input was modified"
    return yd.cpu().numpy()


def c2c(x, dims, direction=FWD, dim_type=S.ASCEND_FFT_HORIZONTAL, precision=0):
    torch = _torch()
    batch = x.size // int(np.prod(dims))
    h = make(dims, S.ASCEND_FFT_C2C, direction, batch, dim_type, precision)
    try:
        return run(h, "c2c", x, x.shape, torch.complex64)
    finally:
        h.destroy()


def r2c(x, dims, direction=FWD):
    torch = _torch()
    batch = x.size // int(np.prod(dims))
    out_shape = x.shape[:-1] + (dims[-1] // 2 + 1,)
    h = make(dims, S.ASCEND_FFT_R2C, direction, batch)
    try:
        return run(h, "r2c", x, out_shape, torch.complex64)
    finally:
        h.destroy()


def c2r(x, dims, direction=INV):
    torch = _torch()
    per = int(np.prod(dims[:-1])) * (dims[-1] // 2 + 1)
    batch = x.size // per
    out_shape = x.shape[:-1] + (dims[-1],)
    h = make(dims, S.ASCEND_FFT_C2R, direction, batch)
    try:
        return run(h, "c2r", x, out_shape, torch.float32)
    finally:
        h.destroy()


def report(tag, got, ref):
    rel, mx = oracle.error(got, ref)
    print(f"[parity] {tag}: rel-L2 {rel:.3e} max-abs {mx:.3e}", flush=True)
    return rel, mx
