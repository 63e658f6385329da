"""Host-side tests of the C-ABI library (no GPU needed): exported symbols, status codes of the
validation paths (restated from core/fft/fft_api.cpp:737-1233), the aclTensor stand-in, the plan
layout, the reference's core-selection rules (core/fft/fft_api.cpp:169-245, 383-697,
core/fft/fftcore/select_core.cpp:155-200) against the resolved values of the hot-path
configurations, and a C++ caller compiled against include/fft_api.h exactly as the reference's
examples call it."""
import ctypes
import os
import re
import subprocess
import threading

import numpy as np

import pytest

import oracle
import sip_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INV, OK, INTERNAL, UNSUP = S.ACL_ERROR_INVALID_PARAM, S.ACL_SUCCESS, S.ACL_ERROR_INTERNAL_ERROR, \
    S.ACL_ERROR_API_NOT_SUPPORT
FWD = S.ASCEND_FFT_FORWARD


def header_functions():
    names = set()
    for h in ["include/fft_api.h", "include/sip_fft.h", "include/aclnn/acl_meta.h"]:
        text = open(os.path.join(ROOT, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)   # drop comments
        text = re.sub(r"//[^\n]*", "", text)
        names |= set(re.findall(r"\b((?:asdFft|acl|sipFft)[A-Za-z0-9]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = S.lib()
    names = header_functions()
    assert {"asdFftCreate", "asdFftExecC2C", "aclCreateTensor", "sipFftPlanDescribeDry"} <= names
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(S.EXPORTED_SYMBOLS) <= names


def test_create_destroy_and_bad_handles():
    L = S.lib()
    h = ctypes.c_void_p()
    assert L.asdFftCreate(ctypes.byref(h)) == OK
    t = ctypes.c_int32()
    assert L.asdFftGetType(h, ctypes.byref(t)) == OK and t.value == S.ASCEND_FFT_C2C
    assert L.asdFftDestroy(h) == OK
    assert L.asdFftDestroy(h) == INV                       # double destroy
    bogus = ctypes.c_void_p(0x1234)
    assert L.asdFftMakePlan1D(bogus, 1024, S.ASCEND_FFT_C2C, FWD, 1, S.ASCEND_FFT_HORIZONTAL) == INV
    assert L.asdFftSetWorkspace(bogus, None) == INV
    assert L.asdFftSynchronize(bogus) == INTERNAL          # fft_api.cpp:963-973
    assert L.asdFftGetType(bogus, ctypes.byref(t)) == INTERNAL   # fft_api.cpp:975-987
    assert L.asdFftCreate(None) == INV


def test_plan_parameter_validation():
    L = S.lib()
    h = S.FftHandle()
    for n in (0, -1, (1 << 27) + 1):
        assert h.make_plan_1d(n, S.ASCEND_FFT_C2C, FWD, 1) == INV
    assert h.make_plan_1d(1024, S.ASCEND_FFT_C2C, FWD, 0) == INV
    assert h.make_plan_1d(1024, 0x99, FWD, 1) == INV
    assert h.make_plan_2d(0, 16, S.ASCEND_FFT_C2C, FWD, 1) == INV
    assert h.make_plan_2d(16, 1 << 28, S.ASCEND_FFT_C2C, FWD, 1) == INV
    assert h.make_plan_2d(16, 16, S.ASCEND_FFT_C2C, FWD, 0) == INV
    assert h.make_plan_3d(16, 16, 0, S.ASCEND_FFT_C2C, FWD, 1) == INV
    assert L.asdFftSetStream(h.h, None) == INV            # fft_api.cpp:747-750
    size = ctypes.c_size_t(7)
    assert L.asdFftGetWorkspaceSize(h.h, ctypes.byref(size)) == OK and size.value == 0
    h.destroy()


def test_exec_on_uninitialised_plan_and_bad_tensors():
    L = S.lib()
    h = S.FftHandle()
    one = S.AclTensor(0x1000, [1], S.ACL_COMPLEX64)
    assert L.asdFftExecC2C(h.h, ctypes.c_void_p(one.ptr), ctypes.c_void_p(one.ptr)) == INV
    assert L.asdFftExecC2C(h.h, None, None) == INV
    # separated exec on a plan that is not C2C_SEP: invalid parameter; ISTFT answers "not supported"
    assert L.asdFftExecC2CSeparated(h.h, one.ptr, one.ptr, one.ptr, one.ptr) == INV
    # ISTFT: a 1-D input is not a (channel, n_fft, frames) spectrum; exec without an ISTFT plan
    assert L.asdFftIstftMakePlan(h.h, one.ptr, 16, 4, 16, True, False, False, 0, True) == INV
    assert L.asdFftExecIstft(h.h, one.ptr, None, one.ptr) == INV
    one.close()
    h.destroy()


def test_istft_make_plan_validation():
    """IstftParamsCheck rules (core/fft/istft.cpp:305-410), all answered before any device work."""
    L = S.lib()
    h = S.FftHandle()

    def mk(shape, n_fft, hop, win, handle=None):
        t = S.AclTensor(0x1000, shape, S.ACL_COMPLEX64)
        try:
            return L.asdFftIstftMakePlan(h.h if handle is None else handle, t.ptr, n_fft, hop, win, True, False,
                                         False, 0, True)
        finally:
            t.close()

    assert mk([2, 16, 1], 16, 4, 16) == INV            # frames must exceed 1
    assert mk([2, 16, 5], 32, 4, 32) == INV            # n_fft must equal the frequency dim
    assert mk([2, 16, 5], 16, 0, 16) == INV            # 0 < hop
    assert mk([2, 16, 5], 16, 17, 16) == INV           # hop <= win_length
    assert mk([2, 16, 5], 16, 4, 12) == INV            # win_length == n_fft
    assert mk([1, 1500, 3], 1500, 4, 1500) == INV      # n_fft < 1500
    assert mk([2, 16, 5], 16, 4, 16, handle=ctypes.c_void_p(0x12345)) == INV  # unknown handle
    real = S.AclTensor(0x1000, [2, 16, 5], S.ACL_FLOAT)
    assert L.asdFftIstftMakePlan(h.h, real.ptr, 16, 4, 16, True, False, False, 0, True) == INV  # complex64 only
    real.close()
    h.destroy()


def test_istft_make_plan_failure_resets_the_handle():
    """A valid ISTFT request whose plan cannot be built (here: no GPU to hold its tables) returns
    500000, and the handle keeps no ISTFT description, so exec answers 100000."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present: the plan builds")
    except ImportError:
        pass
    L = S.lib()
    h = S.FftHandle()
    x = S.AclTensor(0x1000, [2, 16, 5], S.ACL_COMPLEX64)
    w = S.AclTensor(0x2000, [16], S.ACL_FLOAT)
    y = S.AclTensor(0x3000, [2, 16 + 4 * 4 - 16], S.ACL_COMPLEX64)
    assert L.asdFftIstftMakePlan(h.h, x.ptr, 16, 4, 16, True, False, False, 0, True) == INTERNAL
    assert L.asdFftExecIstft(h.h, x.ptr, w.ptr, y.ptr) == INV
    for t in (x, w, y):
        t.close()
    h.destroy()


def test_acl_tensor_descriptor():
    L = S.lib()
    t = S.AclTensor(0x2000, [3, 5, 7], S.ACL_FLOAT)
    dt = ctypes.c_int()
    assert L.aclGetDataType(t.ptr, ctypes.byref(dt)) == 0 and dt.value == S.ACL_FLOAT
    dims = ctypes.POINTER(ctypes.c_int64)()
    nd = ctypes.c_uint64()
    assert L.aclGetViewShape(t.ptr, ctypes.byref(dims), ctypes.byref(nd)) == 0
    assert nd.value == 3 and [dims[i] for i in range(3)] == [3, 5, 7]
    L.aclFreeViewShape(dims)
    assert L.aclGetTensorData(t.ptr) == 0x2000
    t.close()


def dry(dims, t, batch, dim_type=S.ASCEND_FFT_HORIZONTAL, direction=FWD):
    return S.describe_dry(dims, t, direction, batch, dim_type)


@pytest.mark.parametrize("dims,t,batch,expect", [
    ([1024], S.ASCEND_FFT_C2C, 256, ["FftB(N=1024,batch=256,radix=[16,64])"]),
    ([4096], S.ASCEND_FFT_C2C, 16384, ["FftB(N=4096,batch=16384,radix=[64,64])"]),
    ([8192], S.ASCEND_FFT_R2C, 4096, ["FftR2C(N=8192,batch=4096,c2c=4096,radix=[64,64])"]),
    ([8192], S.ASCEND_FFT_C2R, 4096, ["FftC2R(N=8192,batch=4096,c2c=4096,radix=[64,64])"]),
    ([1536], S.ASCEND_FFT_C2C, 8192, ["FftMix(N=1536,batch=8192,radix=[32,48])"]),
    ([2000], S.ASCEND_FFT_C2C, 8192, ["FftMix(N=2000,batch=8192,radix=[40,50])"]),
    ([128], S.ASCEND_FFT_C2C, 32, ["Dft(N=128,batch=32)"]),
    ([32768], S.ASCEND_FFT_C2C, 4, ["FftN(N=32768,batch=4)"]),
    ([512], S.ASCEND_FFT_C2C, 4, ["FftB(N=512,batch=4,radix=[16,32])"]),
    ([256], S.ASCEND_FFT_R2C, 4, ["DftR2C(N=256,batch=4)"]),
    ([199 * 199], S.ASCEND_FFT_C2C, 32, ["Any(N=39601,batch=32)"]),
    ([15000], S.ASCEND_FFT_C2C, 32, None),
    ([1024, 1024], S.ASCEND_FFT_C2C, 8, ["FftB(N=1024,batch=8192,radix=[16,64])", "Transpose(-1,-2,dims=[8,1024,1024])",
                                          "FftB(N=1024,batch=8192,radix=[16,64])", "Transpose(-1,-2,dims=[8,1024,1024])"]),
    ([128, 128, 128], S.ASCEND_FFT_C2C, 4, ["Dd(X=128,Y=128,batch=512)", "Transpose(-1,-3,dims=[4,128,128,128])",
                                             "Dft(N=128,batch=65536)", "Transpose(-1,-3,dims=[4,128,128,128])"]),
    ([64, 64], S.ASCEND_FFT_C2C, 2, ["Dd(X=64,Y=64,batch=2)"]),
])
def test_reference_core_selection(dims, t, batch, expect):
    steps = oracle.reference_steps(dims, t, batch)
    if expect is None:
        assert steps[0].startswith("FftMix(N=15000")
    else:
        assert steps == expect


def test_vertical_reference_rule():
    steps = oracle.reference_steps([4096], S.ASCEND_FFT_C2C, 128, S.ASCEND_FFT_VERTICAL)
    assert steps == ["FftStride(N=4096,batch=128,stride=128)"]


def test_reference_istft_steps():
    assert oracle.reference_istft_steps(2, 400, 11, 100) == [
        "Transpose(1,2,dims=[2,400,11])", "FftMix(N=400,batch=22,radix=[20,20])",
        "IstftAny(n_fft=400,hop=100,frames=11,channel=2)"]


def test_product_library_holds_no_reference_rules():
    """The restated reference rules are test infrastructure: libsip_fft.so's plan description
    carries only this library's own stages."""
    assert "reference_steps" not in dry([4096], S.ASCEND_FFT_C2C, 16384)


def test_plan_layouts():
    d = dry([4096], S.ASCEND_FFT_C2C, 16384)
    assert d["workspace"] == 0 and len(d["stages"]) == 1
    st = d["stages"][0]
    assert st["L"] == 4096 and st["radices"] == [16, 16, 16] and st["direct_in"] == 1 and st["direct_out"] == 1
    d = dry([65536], S.ASCEND_FFT_C2C, 3)
    assert [s["L"] for s in d["stages"]] == [256, 256] and d["stages"][0]["tw4"] == 65536
    assert d["workspace"] >= 65536 * 3 * 8
    d = dry([997], S.ASCEND_FFT_C2C, 2)
    assert [s["kind"] for s in d["stages"]] == ["aux", "stockham", "aux", "stockham", "aux"]
    d = dry([1 << 27], S.ASCEND_FFT_C2C, 1)
    assert len(d["stages"]) == 3 and d["workspace"] == (1 << 27) * 8
    d = dry([1024, 1024], S.ASCEND_FFT_C2C, 8)
    assert d["workspace"] == 0 and len(d["stages"]) == 2 and d["stages"][1]["unit_f_in"] == 1
    d = dry([128, 128, 128], S.ASCEND_FFT_C2C, 4)
    # the last two axes in one fused 2D tile launch, then the first axis in place
    assert d["workspace"] == 0 and len(d["stages"]) == 2 and d["stages"][0]["tile2d"] == [128, 128]
    assert d["stages"][1]["in"] == d["stages"][1]["out"] == "out"
    d = dry([64, 32], S.ASCEND_FFT_C2C, 3)
    assert len(d["stages"]) == 1 and d["stages"][0]["tile2d"] == [64, 32] and d["stages"][0]["grid"] == 3
    d = dry([8192], S.ASCEND_FFT_R2C, 4096)
    assert len(d["stages"]) == 1 and d["stages"][0]["io"] == 1 and d["stages"][0]["L"] == 4096
    d = dry([100, 200], S.ASCEND_FFT_C2R, 2)
    assert d["workspace"] > 0  # the input is not modified, so the complex axis goes to workspace


@pytest.mark.parametrize("n", [16875, 19683, 20000, 24576, 30375, 32768])
def test_single_pass_long_rows(n):
    """2,3,5-smooth rows above 16384 with a radix sequence of at most 32 values per thread, 256 or
    more of them: one single-plane launch of 1024 threads per transform, no workspace; as the last
    axis of a 2D transform as well, and for the half-length transform of an R2C row. Fewer rows keep
    the four-step split."""
    d = dry([n], S.ASCEND_FFT_C2C, 256)
    assert d["workspace"] == 0 and len(d["stages"]) == 1
    st = d["stages"][0]
    assert st["L"] == n and st["F"] == 1 and st["split_lds"] == 1 and st["direct_in"] == st["direct_out"] == 1
    assert int(np.prod(st["radices"])) == n and st["T"] <= 1024 and st["lds"] <= 160 * 1024
    assert len(dry([n], S.ASCEND_FFT_C2C, 255)["stages"]) == 2
    d = dry([6, n], S.ASCEND_FFT_C2C, 50)
    assert d["stages"][0]["L"] == n and d["stages"][0]["nfft"] == 300
    d = dry([2 * n], S.ASCEND_FFT_R2C, 300)
    assert d["stages"][0]["L"] == n and d["stages"][0]["note"] == f"leaf {n}"


@pytest.mark.parametrize("n", [31104, 32000, 65536, 3 ** 10])
def test_long_rows_without_single_pass_sequence(n):
    """Lengths with no sequence within 32 values per thread (31104, 32000) and lengths above 32768
    keep the four-step split."""
    d = dry([n], S.ASCEND_FFT_C2C, 512)
    assert len(d["stages"]) >= 2 and d["stages"][0]["tw4"] == n


def test_streaming_access_choices():
    """Nontemporal stores on a stage whose output no later stage reads (column outputs only from
    512 MiB up); nontemporal loads on C2C stages that read 512 MiB or more."""
    st = dry([4096], S.ASCEND_FFT_C2C, 16384)["stages"][0]          # 512 MiB in
    assert st["nt_out"] == 1 and st["nt_in"] == 1
    st = dry([4096], S.ASCEND_FFT_C2C, 16383)["stages"][0]          # just below
    assert st["nt_out"] == 1 and st["nt_in"] == 0
    d = dry([1024, 1024], S.ASCEND_FFT_C2C, 8)                      # rows then columns in place, 64 MiB
    assert [s["nt_out"] for s in d["stages"]] == [0, 0] and [s["nt_in"] for s in d["stages"]] == [0, 0]
    d = dry([1024, 1024], S.ASCEND_FFT_C2C, 128)                    # 1 GiB
    assert [s["nt_out"] for s in d["stages"]] == [0, 1] and [s["nt_in"] for s in d["stages"]] == [1, 1]
    d = dry([65536], S.ASCEND_FFT_C2C, 2048)                        # four-step, 1 GiB per stage
    assert [s["nt_out"] for s in d["stages"]] == [0, 1] and [s["nt_in"] for s in d["stages"]] == [1, 1]
    st = dry([8192], S.ASCEND_FFT_R2C, 1 << 16)["stages"][0]        # real rows keep cached loads
    assert st["nt_in"] == 0
    st = dry([2000], S.ASCEND_FFT_C2C, 1 << 16)["stages"][0]        # T = 100: runs not whole lines
    assert st["T"] == 100 and st["nt_out"] == 0 and st["nt_in"] == 0
    st = dry([1536], S.ASCEND_FFT_C2C, 1 << 16)["stages"][0]        # T = 96, radix-16 first / radix-6 last
    assert st["T"] == 96 and st["nt_out"] == 1 and st["nt_in"] == 1


def test_single_plane_kernel_choices():
    """The launch shapes the measured defaults rely on: one LDS plane for contiguous C2C rows, the
    fused real-transform rows, column tiles (with the four-step twiddle on store for stage A) and the
    128 x 128 slice; two planes for the four-step row stage of 16 threads per transform."""
    st = dry([4096], S.ASCEND_FFT_C2C, 16384)["stages"][0]
    assert (st["split_lds"], st["lds"], st["block"], st["F"]) == (1, 17404, 256, 1)
    for t in (S.ASCEND_FFT_R2C, S.ASCEND_FFT_C2R):
        st = dry([8192], t, 4096)["stages"][0]
        assert st["split_lds"] == 1 and st["L"] == 4096 and st["lds"] < 20 * 1024
    rows, cols = dry([1024, 1024], S.ASCEND_FFT_C2C, 8)["stages"]
    assert rows["split_lds"] == 1 and cols["split_lds"] == 1 and cols["col_direct"] == 1
    assert cols["F"] == 16 and cols["lds"] <= 80 * 1024  # two 1024-thread tiles per CU
    a, b = dry([65536], S.ASCEND_FFT_C2C, 3)["stages"]
    assert a["split_lds"] == 1 and a["col_direct"] == 1 and a["tw4"] == 65536 and b["split_lds"] == 0
    st = dry([128, 128, 128], S.ASCEND_FFT_C2C, 4)["stages"][0]
    assert st["tile2d"] == [128, 128] and st["lds"] <= 80 * 1024
    cols = dry([1000, 1000], S.ASCEND_FFT_C2C, 2)["stages"][1]
    assert cols["split_lds"] == 1 and cols["F"] == 16


def test_registry_is_thread_safe():
    L = S.lib()
    errors = []

    def worker():
        for _ in range(200):
            h = ctypes.c_void_p()
            if L.asdFftCreate(ctypes.byref(h)) != OK or L.asdFftDestroy(h) != OK:
                errors.append(1)

    th = [threading.Thread(target=worker) for _ in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors


def test_cpp_caller_compiles_against_reference_header(tmp_path):
    """A C++ caller written against include/fft_api.h (references, default argument, AsdSip enums),
    as the reference's examples are, links against libsip_fft.so and runs its host-side checks."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include <cstdio>
#include "fft_api.h"
#include "aclnn/acl_meta.h"
using namespace AsdSip;
int main() {
    asdFftHandle h;
    if (asdFftCreate(h) != ACL_SUCCESS) return 1;
    if (asdFftMakePlan1D(h, 0, ASCEND_FFT_C2C, ASCEND_FFT_FORWARD, 4) != ACL_ERROR_INVALID_PARAM) return 2;
    asdFftType t;
    if (asdFftGetType(h, t) != ACL_SUCCESS || t != ASCEND_FFT_C2C) return 3;
    size_t ws = 99;
    if (asdFftGetWorkspaceSize(h, ws) != ACL_SUCCESS || ws != 0) return 4;
    int64_t dims[2] = {4, 16};
    aclTensor *x = aclCreateTensor(dims, 2, ACL_COMPLEX64, nullptr, 0, ACL_FORMAT_ND, dims, 2, nullptr);
    if (!x) return 5;
    aclDataType dt;
    if (aclGetDataType(x, &dt) != 0 || dt != ACL_COMPLEX64) return 6;
    aclDestroyTensor(x);
    if (asdFftDestroy(h) != ACL_SUCCESS) return 7;
    std::puts("ok");
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(S.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L" + libdir, "-lsip_fft", "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "ok" in r.stdout, (r.returncode, r.stdout, r.stderr)


def test_set_precision_validation():
    """sipFftSetPrecision accepts modes 0..3 only and rejects unknown handles with 100000."""
    h = S.FftHandle()
    try:
        assert h.set_precision(7) == S.ACL_ERROR_INVALID_PARAM
        assert h.set_precision(-1) == S.ACL_ERROR_INVALID_PARAM
    finally:
        h.destroy()
    assert S.lib().sipFftSetPrecision(None, 1) == S.ACL_ERROR_INVALID_PARAM


def _seq_override():
    import importlib.util
    path = os.path.join(ROOT, "sip_amd", "csrc", "kernels", "gen_instances.py")
    spec = importlib.util.spec_from_file_location("gen_instances_for_test", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.SEQ_OVERRIDE


def test_measured_radix_sequences_are_built():
    """Every length whose radix sequence was chosen by measurement runs that sequence (as a leaf of
    its own length: from 256 rows up for the lengths above 16384)."""
    for n, seq in sorted(_seq_override().items()):
        st = dry([n], S.ASCEND_FFT_C2C, 256)["stages"][0]
        assert st["L"] == n and st["radices"] == list(seq), (n, st["radices"], seq)
